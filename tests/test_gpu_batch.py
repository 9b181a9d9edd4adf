"""Batched ensembles: a scenario's state is bitwise the same alone and in a
batch of uneven grids with mixed tile heights; the autotuner leaves the state
untouched; graph replay equals step-by-step launches."""
import sys

import numpy as np
import pytest

from conftest import ensure_oracle, random_case

pytestmark = pytest.mark.gpu
ensure_oracle()
import oracle_py as O  # noqa: E402
import sw2d  # noqa: E402


def say(m):
    print(m, flush=True)
    sys.stdout.flush()


SHAPES = [(70, 130), (256, 256), (33, 400), (129, 65), (1, 90), (200, 150)]


def scen_inputs(k, shape):
    z, h = random_case(500 + k, *shape, holes=0.02)
    rng = np.random.default_rng(600 + k)
    return z, h, rng.uniform(0.02, 0.1, shape), rng.gamma(0.8, 20.0, 24) * 300 / 3600.0


@pytest.mark.parametrize("prec", [sw2d.FP64, sw2d.FP32])
def test_batch_equals_alone(prec):
    say("case: batch of %d uneven grids vs alone, precision %d" % (len(SHAPES), prec))
    kw = dict(dt=0.5, manning=0.03, cfl=0.45, edge_open=sw2d.EDGE_SOUTH, forcing_mode=sw2d.FORCING_STEP,
              INT=0.3, INF=0.1, LWL=0.2, EV_WL_min=0.05, t_end=150.0, precision=prec)
    b = sw2d.Batch(prec)
    ids = []
    for k, shp in enumerate(SHAPES):
        z, h, n, rain = scen_inputs(k, shp)
        sid = b.add(z, h, 2.0, n_cell=n, **kw)
        b.scenario(sid).set_rain(rain, 300.0)
        ids.append(sid)
    rows = [4, 8, 16 if prec == sw2d.FP32 else 8, 4, 8, 0]
    for sid, r in zip(ids, rows):
        b.set_tile_rows(sid, r)
    b.touch()
    for _ in range(20):
        b.step(50)
    done, iters = b.status()
    say("done %s iters %s" % (done.tolist(), iters.tolist()))
    assert done.all()
    for k, shp in enumerate(SHAPES):
        z, h, n, rain = scen_inputs(k, shp)
        a = sw2d.Sim(z, h, 2.0, n_cell=n, **kw)
        a.set_rain(rain, 300.0)
        a.step(int(iters[k]) + 10)
        sa = a.stats()
        s = b.scenario(ids[k])
        sb = s.stats()
        assert sa["iterations"] == sb["iterations"] == iters[k]
        assert sa["t"] == sb["t"]
        for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V):
            x, y = a.read(f), s.read(f)
            assert np.array_equal(x, y), (k, f, float(np.abs(x - y).max()))
        assert abs(sa["rain_in"] - sb["rain_in"]) <= 1e-10 * max(1.0, sa["rain_in"])


def test_batch_fp64_scenario_matches_oracle():
    say("case: batch FP64 scenario vs oracle")
    b = sw2d.Batch(sw2d.FP64)
    z, h, n, rain = scen_inputs(1, (96, 130))
    kw = dict(dt=0.05, manning=0.03, edge_open=sw2d.EDGE_EAST)
    other = b.add(*random_case(9, 50, 70)[:2], 2.0, dt=0.05, manning=0.03)
    sid = b.add(z, h, 2.0, n_cell=n, **kw)
    b.set_tile_rows(sid, 4)
    b.scenario(sid).set_rain(rain, 300.0)
    b.touch()
    b.step(200)
    o = O.OracleSim(z, (z != -9999.0).astype(np.uint8), h, 2.0, 0.05, 0.03, n_cell=n,
                    edge_open=O.EDGE_EAST)
    o.set_rain(rain, 300.0)
    o.step(200)
    s = b.scenario(sid)
    for fo, fg in [(O.H, sw2d.F_H), (O.U, sw2d.F_U), (O.V, sw2d.F_V)]:
        assert np.abs(o.read(fo) - s.read(fg)).max() <= 1e-10
    assert b.status()[1][other] == 200


def test_autotune_keeps_state_and_picks_valid_rows():
    say("case: autotune")
    b = sw2d.Batch(sw2d.FP32)
    ids = [b.add(*random_case(700 + k, *shp)[:2], 2.0, dt=0.05, manning=0.03, precision=sw2d.FP32)
           for k, shp in enumerate([(300, 300), (64, 64), (1024, 200)])]
    b.step(10)
    before = [b.scenario(i).read(sw2d.F_H) for i in ids]
    b.autotune(5)
    after = [b.scenario(i).read(sw2d.F_H) for i in ids]
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    assert all(b.tile_rows(i) % 4 == 0 and b.tile_rows(i) >= 4 for i in ids)
    b.step(10)
    assert (b.status()[1] == 20).all()


def test_shared_statics_and_remove():
    say("case: shared statics + remove")
    b = sw2d.Batch(sw2d.FP32)
    z, h, n, rain = scen_inputs(3, (128, 192))
    kw = dict(dt=0.2, manning=0.03, cfl=0.4, precision=sw2d.FP32, forcing_mode=sw2d.FORCING_STEP)
    a = b.add(z, h, 2.0, n_cell=n, **kw)
    c = b.add(z, h, 2.0, n_cell=n, share_with=a, **kw)
    b.scenario(a).set_rain(rain, 300.0)
    b.scenario(c).set_rain(rain * 2.0, 300.0)
    b.touch()
    b.step(100)
    alone = sw2d.Sim(z, h, 2.0, n_cell=n, **kw)
    alone.set_rain(rain * 2.0, 300.0)
    alone.step(100)
    assert np.array_equal(alone.read(sw2d.F_H), b.scenario(c).read(sw2d.F_H))
    with pytest.raises(sw2d.SW2DError):
        b.remove(a)     # c still shares its statics
    b.remove(c)
    b.step(10)
    assert b.status()[1][a] == 110


def test_graph_replay_equals_single_launches():
    say("case: graph replay vs single launches")
    z, h = random_case(77, 90, 140)
    a = sw2d.Sim(z, h, 2.0, dt=0.05, manning=0.03, precision=sw2d.FP32, cfl=0.5)
    b = sw2d.Sim(z, h, 2.0, dt=0.05, manning=0.03, precision=sw2d.FP32, cfl=0.5)
    a.step(160)              # 5 graph replays
    for _ in range(160):
        b.step(1)            # no graph
    for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V):
        assert np.array_equal(a.read(f), b.read(f))
    assert a.stats()["t"] == b.stats()["t"]


def test_rebuild_between_steps_without_sync():
    """Changing a scenario's segment height while iterations are still in
    flight (no sync between the steps): the rebuild of the work-list waits
    for the batch stream, so the state equals an alone run bitwise."""
    say("case: set_tile_rows between asynchronous batch steps")
    kw = dict(dt=0.2, manning=0.03, cfl=0.45, precision=sw2d.FP32, forcing_mode=sw2d.FORCING_STEP)
    b = sw2d.Batch(sw2d.FP32)
    shapes = [(300, 260), (180, 400)]
    ids = []
    for k, shp in enumerate(shapes):
        z, h, n, rain = scen_inputs(20 + k, shp)
        sid = b.add(z, h, 2.0, n_cell=n, **kw)
        b.scenario(sid).set_rain(rain, 300.0)
        ids.append(sid)
    b.touch()
    b.step(150)                      # asynchronous: graph replays in flight
    b.set_tile_rows(ids[0], 8)
    b.set_tile_rows(ids[1], 24)
    b.step(90)
    b.set_tile_rows(ids[0], 0)
    b.step(37)
    for k, shp in enumerate(shapes):
        z, h, n, rain = scen_inputs(20 + k, shp)
        a = sw2d.Sim(z, h, 2.0, n_cell=n, **kw)
        a.set_rain(rain, 300.0)
        a.step(277)
        s = b.scenario(ids[k])
        assert s.stats()["iterations"] == 277
        for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V):
            assert np.array_equal(a.read(f), s.read(f)), (k, f)


def test_batch_scenario_context_is_not_stepped_alone():
    say("case: sw2d_step on a batch scenario is refused")
    b = sw2d.Batch(sw2d.FP32)
    sid = b.add(*random_case(31, 40, 60)[:2], 2.0, dt=0.05, manning=0.03, precision=sw2d.FP32)
    with pytest.raises(sw2d.SW2DError) as e:
        b.scenario(sid).step(1)
    assert e.value.code == sw2d.E_STATE
    b.step(3)
    assert b.status()[1][sid] == 3


def test_detach_continues_alone_bitwise():
    """A scenario taken out of the batch with sw2d_batch_detach while
    iterations are in flight keeps its state, clock and accumulators: stepped
    on its own afterwards it equals a run that was alone throughout, bitwise;
    the scenarios left in the batch go on; a scenario whose statics another
    shares is refused."""
    say("case: detach a batch scenario mid-run, continue it alone")
    kw = dict(dt=0.2, manning=0.03, cfl=0.45, precision=sw2d.FP32, forcing_mode=sw2d.FORCING_STEP,
              edge_open=sw2d.EDGE_SOUTH)
    b = sw2d.Batch(sw2d.FP32)
    shapes = [(160, 230), (210, 120)]
    ids = []
    for k, shp in enumerate(shapes):
        z, h, n, rain = scen_inputs(40 + k, shp)
        sid = b.add(z, h, 2.0, n_cell=n, **kw)
        b.scenario(sid).set_rain(rain, 300.0)
        ids.append(sid)
    z0, h0, n0, _ = scen_inputs(40, shapes[0])
    sib = b.add(z0, h0, 2.0, n_cell=n0, share_with=ids[0], **kw)
    b.touch()
    b.step(120)                      # asynchronous: graph replays in flight
    with pytest.raises(sw2d.SW2DError) as e:
        b.detach(ids[0])             # its statics are shared by `sib`
    assert e.value.code == sw2d.E_STATE
    d = b.detach(ids[1])
    d.step(80)
    b.step(30)
    z, h, n, rain = scen_inputs(41, shapes[1])
    a = sw2d.Sim(z, h, 2.0, n_cell=n, **kw)
    a.set_rain(rain, 300.0)
    a.step(200)
    sa, sd = a.stats(), d.stats()
    assert sa["iterations"] == sd["iterations"] == 200 and sa["t"] == sd["t"]
    for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V):
        assert np.array_equal(a.read(f), d.read(f)), f
    # volume sums: per-segment partial sums, so the segment heights (batch
    # and alone differ) change their last bits only
    for k in ("rain_in", "outflow"):
        assert abs(sa[k] - sd[k]) <= 1e-10 * max(1.0, abs(sa[k])), k
    d.close()
    assert b.status()[1][ids[0]] == 150 and b.status()[1][sib] == 150


def test_attach_a_stepped_context():
    """A context stepped on its own, then attached to a running batch with
    sw2d_batch_attach, continues from its state: after more batch steps it
    equals a run that was alone throughout, bitwise; a context already in a
    batch is refused; detached and attached again, it takes its old id and
    still matches."""
    say("case: attach a context stepped alone to a running batch")
    kw = dict(dt=0.2, manning=0.03, cfl=0.45, precision=sw2d.FP32, forcing_mode=sw2d.FORCING_STEP)
    b = sw2d.Batch(sw2d.FP32)
    z, h, n, rain = scen_inputs(50, (150, 170))
    first = b.add(z, h, 2.0, n_cell=n, **kw)
    b.scenario(first).set_rain(rain, 300.0)
    b.touch()
    b.step(64)                       # asynchronous: in flight during the attach
    z2, h2, n2, rain2 = scen_inputs(51, (190, 110))
    c = sw2d.Sim(z2, h2, 2.0, n_cell=n2, **kw)
    c.set_rain(rain2, 300.0)
    c.step(45)
    sid = b.attach(c)
    with pytest.raises(sw2d.SW2DError) as e:
        b.attach(b.scenario(first))  # already a batch scenario
    assert e.value.code == sw2d.E_STATE
    b.step(70)
    a = sw2d.Sim(z2, h2, 2.0, n_cell=n2, **kw)
    a.set_rain(rain2, 300.0)
    a.step(115)
    s = b.scenario(sid)
    assert s.stats()["iterations"] == 115 and s.stats()["t"] == a.stats()["t"]
    for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V):
        assert np.array_equal(a.read(f), s.read(f)), f
    assert b.status()[1][first] == 134
    # the id a detach frees is given to the next attach, and the scenario
    # taken out and put back goes on from where it was
    back = b.detach(sid)
    assert b.attach(back) == sid
    b.step(20)
    a.step(20)
    s = b.scenario(sid)
    assert s.stats()["iterations"] == 135
    for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V):
        assert np.array_equal(a.read(f), s.read(f)), f
    assert b.status()[1][first] == 154
