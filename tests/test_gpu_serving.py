"""Serving loop (sw2d_server_*): continuous batching of requests that
arrive while others run.  Each result must be the one the same scenario
gives alone (sw2d_create + sw2d_step), bitwise in h, U and V; admission
follows arrival order within the memory budget, activation least estimated
work first."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "sw2d-gpu_amd"))
import sw2d  # noqa: E402


def say(msg):
    print(msg, flush=True)


def request(k):
    """Small seeded catchment requests of uneven size and duration."""
    side = [40, 72, 56, 96, 48, 80][k % 6]
    z = sw2d.gen_fractal_dem(side, side, 3000 + k, 0.6, 0.0, 8.0)
    z = sw2d.gen_tilt(z, 5.0, 0.005, sw2d.EDGE_SOUTH)
    n = sw2d.gen_uniform((side, side), 3100 + k, 0.02, 0.12)
    h0 = np.zeros_like(z)
    h0[side // 3: side // 2, side // 3: side // 2] = 0.5
    rain = sw2d.gen_gamma((24,), 3200 + k, 0.8, 20.0) * (300.0 / 3600.0)
    params = dict(dt=1.0, manning=0.05, cfl=0.5, edge_open=sw2d.EDGE_SOUTH,
                  forcing_mode=sw2d.FORCING_STEP, t_end=60.0 + 35.0 * k)
    return z, h0, n, rain, params


def run_alone(precision, z, h0, n, rain, params):
    s = sw2d.Sim(z, h0, 5.0, n_cell=n, precision=precision, **params)
    s.set_rain(rain, 300.0)
    while not s.stats()["done"]:
        s.step(50)
    return s


@pytest.mark.parametrize("precision", [sw2d.FP32, sw2d.FP64])
def test_served_results_equal_alone(precision):
    say("case: 6 requests arriving in 3 waves, precision %d" % precision)
    srv = sw2d.Server(precision, mem_budget=1 << 30)
    ids, waves = [], {0: [0, 1], 2: [2, 3], 4: [4, 5]}
    t = 0
    while True:
        for k in waves.get(t, []):
            z, h0, n, rain, params = request(k)
            ids.append(srv.submit(z, h0, 5.0, n_cell=n, rain=rain, **params))
        running, queued = srv.advance(20)
        t += 1
        if t > 4 and running == 0 and queued == 0:
            break
        assert t < 1000
    for k, rid in enumerate(ids):
        info = srv.info(rid)
        assert info["state"] == sw2d.REQ_DONE
        assert info["submit_s"] <= info["admit_s"] <= info["start_s"] <= info["finish_s"]
        z, h0, n, rain, params = request(k)
        a = run_alone(precision, z, h0, n, rain, params)
        sa, sb = a.stats(), srv.stats(rid)
        say("request %d: %dx%d, %d iterations, t = %.3f" % (k, z.shape[0], z.shape[1],
                                                          info["iterations"], sb["t"]))
        assert sa["iterations"] == sb["iterations"] == info["iterations"]
        assert sa["t"] == sb["t"]
        for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V):
            x, y = a.read(f), srv.read(rid, f)
            assert np.array_equal(x, y), (k, f, float(np.abs(x - y).max()))
        assert abs(sa["volume"] - sb["volume"]) <= 1e-12 * max(1.0, sa["volume"])
        srv.release(rid)
        with pytest.raises(sw2d.SW2DError):
            srv.read(rid, sw2d.F_H)
    srv.close()


def test_admission_follows_arrival_order_within_budget():
    say("case: budget for one request at a time")
    reqs = [request(k) for k in range(4)]
    need = max(sw2d.footprint_bytes(r[0].shape[0], r[0].shape[1], sw2d.FP32, True) for r in reqs)
    srv = sw2d.Server(sw2d.FP32, mem_budget=int(need * 1.05))  # no two requests fit together
    ids = [srv.submit(z, h0, 5.0, n_cell=n, rain=rain, **p) for z, h0, n, rain, p in reqs]
    peak = 0
    for _ in range(2000):
        running, queued = srv.advance(25)
        peak = max(peak, running)
        if running == 0 and queued == 0:
            break
    assert peak == 1
    infos = [srv.info(r) for r in ids]
    assert all(i["state"] == sw2d.REQ_DONE for i in infos)
    admits = [i["admit_s"] for i in infos]
    assert admits == sorted(admits)
    for a, b in zip(infos, infos[1:]):  # each admitted only after the previous finished
        assert b["admit_s"] >= a["finish_s"]
    srv.close()


def test_activation_least_work_first():
    """With room for one active request (an activation target of 1 cell),
    requests admitted together start least estimated work first (cells x
    simulated seconds), whatever their arrival order, and each still gives
    the result of its alone run."""
    say("case: activation order by estimated work, one active at a time")
    srv = sw2d.Server(sw2d.FP32, mem_budget=1 << 30, active_cells=1)
    ks = [3, 0, 5, 1]  # arrival order
    reqs = {k: request(k) for k in ks}
    ids = {k: srv.submit(z, h0, 5.0, n_cell=n, rain=rain, **p) for k, (z, h0, n, rain, p) in reqs.items()}
    peak = 0
    for _ in range(4000):
        running, queued = srv.advance(25)
        peak = max(peak, running)
        if running == 0 and queued == 0:
            break
    assert peak == 4  # all four admitted (resident) together
    infos = {k: srv.info(ids[k]) for k in ks}
    assert all(i["state"] == sw2d.REQ_DONE for i in infos.values())
    work = {k: reqs[k][0].size * reqs[k][4]["t_end"] for k in ks}
    by_start = sorted(ks, key=lambda k: infos[k]["start_s"])
    assert by_start == sorted(ks, key=lambda k: work[k]), (by_start, work)
    for k in ks:
        assert infos[k]["submit_s"] <= infos[k]["admit_s"] <= infos[k]["start_s"] <= infos[k]["finish_s"]
    z, h0, n, rain, p = reqs[3]
    a = run_alone(sw2d.FP32, z, h0, n, rain, p)
    for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V):
        assert np.array_equal(a.read(f), srv.read(ids[3], f)), f
    srv.close()


def test_oversized_request_is_rejected():
    say("case: request larger than the budget")
    srv = sw2d.Server(sw2d.FP32, mem_budget=1 << 17)  # 128 KiB: below any request's footprint
    z, h0, n, rain, p = request(3)
    with pytest.raises(sw2d.SW2DError) as e:
        srv.submit(z, h0, 5.0, n_cell=n, rain=rain, **p)
    assert e.value.code == -4
    srv.close()
