"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

FP64 gate: max |dh|, |dU|, |dV| <= 1e-10 after the stated iterations
(differences come only from last-bit pow differences
between the device and host math libraries).
"""
import gzip
import os
import shutil
import sys

import numpy as np
import pytest

from conftest import DAM_BREAK, dam_break_1d, ensure_oracle, random_case

pytestmark = pytest.mark.gpu

ensure_oracle()
import caseio  # noqa: E402
import oracle_py as O  # noqa: E402
import sw2d  # noqa: E402

TOL64 = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pair(z, h0, dx, dt, manning, *, n_cell=None, precision=sw2d.FP64, **kw):
    act = (z != -9999.0).astype(np.uint8)
    o = O.OracleSim(z, act, h0, dx, dt, manning, n_cell=n_cell, **kw)
    g = sw2d.Sim(z, h0, dx, nodata=-9999.0, n_cell=n_cell, dt=dt, manning=manning,
                 precision=precision, **kw)
    return o, g


def diffs(o, g):
    out = {}
    for name, fo, fg in [("h", O.H, sw2d.F_H), ("U", O.U, sw2d.F_U), ("V", O.V, sw2d.F_V),
                         ("uu", O.UU, sw2d.F_UU), ("vv", O.VV, sw2d.F_VV)]:
        a, b = o.read(fo), g.read(fg)
        assert a.shape == b.shape, name
        out[name] = float(np.max(np.abs(a - b))) if a.size else 0.0
        out[name + "_neq"] = int(np.count_nonzero(a != b))
    return out


def test_gpu_kat_results0_vtk(tmp_path):
    """One iteration on db/dam_break through the product reader, kernels and
    writer reproduces the reference's committed Results_0.vtk."""
    say("case: dam_break KAT, 1 iteration FP64")
    c = sw2d.Case(DAM_BREAK)
    s = c.create()
    s.step(1)
    s.sync()
    out = str(tmp_path / "Results_0.vtk")
    c.write_vtk(out, c.array(sw2d.CASE_IC).reshape(100, 100), s.read(sw2d.F_VV), s.read(sw2d.F_UU),
                c.array(sw2d.CASE_DEM).reshape(100, 100))
    ref = gzip.open(os.path.join(DAM_BREAK, "output", "Results_0.vtk.gz")).read()
    assert open(out, "rb").read() == ref


@pytest.mark.parametrize("transpose", [False, True])
def test_gpu_config1_dam_break_fp64(transpose):
    """Config 1: 200 x 100 dam break, dx = 1, n = 0.01, dt = 0.01, 1000
    iterations, closed walls, no rain."""
    say("case: config1 dam break transpose=%s, 1000 iterations FP64" % transpose)
    z, h = dam_break_1d(transpose=transpose)
    o, g = pair(z, h, 1.0, 0.01, 0.01)
    o.step(1000)
    g.step(1000)
    g.sync()
    d = diffs(o, g)
    say(str(d))
    for k in ("h", "U", "V", "uu", "vv"):
        assert d[k] <= TOL64, d
    st = g.stats()
    assert st["iterations"] == 1000 and abs(st["t"] - o.stats()["t"]) == 0


def test_gpu_config1_rows_identical():
    say("case: config1 rows identical, 300 iterations FP64")
    z, h = dam_break_1d()
    g = sw2d.Sim(z, h, 1.0, dt=0.01, manning=0.01)
    g.step(300)
    H = g.read(sw2d.F_H)
    assert (H == H[0:1]).all()


@pytest.mark.parametrize("seed,shape,edges", [
    (21, (37, 70), 0),
    (22, (64, 129), sw2d.EDGE_SOUTH),
    (23, (19, 200), sw2d.EDGE_NORTH | sw2d.EDGE_EAST),
    (24, (131, 65), sw2d.EDGE_WEST | sw2d.EDGE_SOUTH),
    (25, (1, 77), 0),
    (26, (90, 1), sw2d.EDGE_SOUTH),
])
def test_gpu_random_grids_fp64(seed, shape, edges):
    """Masked random beds, rain (reference formula), a point source, open
    edges and a Manning raster; 300 iterations."""
    say("case: random grid seed=%d shape=%s edges=%d FP64" % (seed, shape, edges))
    z, h = random_case(seed, *shape)
    rng = np.random.default_rng(seed + 100)
    ncell = rng.uniform(0.02, 0.1, shape)
    kw = dict(INT=0.45, INF=0.2, LWL=0.3, EV_WL_min=0.05, edge_open=edges)
    o, g = pair(z, h, 2.0, 0.02, 0.03, n_cell=ncell, **kw)
    table = rng.uniform(0, 20, 30)
    o.set_rain(table, 60.0)
    g.set_rain(table, 60.0)
    q = rng.uniform(-0.5, 2.0, (12, 2))
    cells = [int(rng.integers(0, shape[0] * shape[1])), int(rng.integers(0, shape[0] * shape[1]))]
    o.set_sources(q, 60.0, cells)
    g.set_sources(q, 60.0, cells)
    o.step(300)
    g.step(300)
    g.sync()
    d = diffs(o, g)
    say(str(d))
    for k in ("h", "U", "V", "uu", "vv"):
        assert d[k] <= TOL64, d
    so, sg = o.stats(), g.stats()
    for k in ("volume", "rain_in", "source_in", "outflow", "clamp_gain"):
        assert abs(so[k] - sg[k]) <= 1e-9 * max(1.0, abs(so[k])), (k, so[k], sg[k])


def test_gpu_rain_map_and_evaporation_fp64():
    say("case: rain block map + evaporation FP64")
    z, h = random_case(31, 96, 160, holes=0.02)
    rng = np.random.default_rng(3)
    kw = dict(INT=0.45, INF=0.2, LWL=0.3, EV_WL_min=0.05, evaporation_on=1)
    o, g = pair(z, h, 5.0, 0.05, 0.04, **kw)
    bmap = rng.lognormal(0, 0.5, (6, 10))
    table = rng.uniform(0, 10, 50)
    evap = sw2d.evaporation_table(0.25, rng.uniform(15, 30, 50), rng.uniform(0, 800, 50), 60.0)
    o.set_rain(table, 60.0, bmap, 16)
    g.set_rain(table, 60.0, bmap, 16)
    o.set_evaporation(evap)
    g.set_evaporation(evap)
    o.step(400)
    g.step(400)
    d = diffs(o, g)
    say(str(d))
    for k in ("h", "U", "V"):
        assert d[k] <= TOL64, d


def test_gpu_cfl_fp64():
    """Adaptive step (Courant 0.5, capped at 2*dt): same step sequence."""
    say("case: CFL FP64")
    z, h = random_case(41, 70, 90, holes=0.0, wet=1.0)
    o, g = pair(z, h, 2.0, 0.5, 0.03, cfl=0.5, edge_open=sw2d.EDGE_SOUTH)
    g.record_dt(500)
    o.step(200)
    g.step(200)
    d = diffs(o, g)
    say(str(d))
    for k in ("h", "U", "V"):
        assert d[k] <= TOL64, d
    hist = g.dt_history(500)
    assert len(hist) == 200
    so, sg = o.stats(), g.stats()
    assert abs(so["t"] - sg["t"]) <= 1e-9 and abs(so["last_D"] - sg["last_D"]) <= 1e-12
    assert hist.min() < 1.0   # the step was limited


def test_gpu_dam_break_case_fp64(dam_case):
    say("case: db/dam_break 500 iterations FP64")
    c = dam_case
    d, s = c["dem"], c["setup"]
    kw = dict(INT=s["INT"], INF=s["INF"], LWL=s["LWL"], EV_WL_min=s["EV_WL_min"])
    o, g = pair(d["z"], c["ic"], d["res"], s["dt"], s["manning_coef"], **kw)
    o.step(500)
    g.step(500)
    dd = diffs(o, g)
    say(str(dd))
    for k in ("h", "U", "V", "uu", "vv"):
        assert dd[k] <= TOL64, dd


def test_gpu_snapshot_restore_bitwise():
    say("case: snapshot/restore FP64 + FP32")
    for prec in (sw2d.FP64, sw2d.FP32):
        z, h = random_case(51, 40, 70)
        g = sw2d.Sim(z, h, 2.0, dt=0.05, manning=0.03, precision=prec, cfl=0.4)
        g.set_rain(np.full(20, 5.0), 60.0)
        g.step(30)
        blob = g.snapshot()
        g.step(40)
        a = [g.read(f) for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V)]
        ta = g.stats()
        g.restore(blob)
        g.step(40)
        b = [g.read(f) for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V)]
        tb = g.stats()
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert ta["t"] == tb["t"] and ta["iterations"] == tb["iterations"]


def test_gpu_fp32_against_fp64():
    """FP32 state with the FP64 step sequence replayed: depth within 1e-4 m
    and volume within 1e-6 relative of the FP64 run."""
    say("case: FP32 vs FP64, catchment 256x256, 600 iterations")
    rng = np.random.default_rng(61)
    R = Cc = 256
    ii, jj = np.meshgrid(np.arange(R), np.arange(Cc), indexing="ij")
    z = 40.0 + 0.01 * 5.0 * (R - ii) + rng.uniform(0, 0.5, (R, Cc))
    h = np.zeros((R, Cc))
    ncell = rng.uniform(0.03, 0.10, (R, Cc))
    table = np.full(40, 50.0 * 300 / 3600.0)
    a = sw2d.Sim(z, h, 5.0, dt=1.0, manning=0.05, n_cell=ncell, cfl=0.5, edge_open=sw2d.EDGE_SOUTH,
                 forcing_mode=sw2d.FORCING_STEP, precision=sw2d.FP64)
    a.set_rain(table, 300.0)
    a.record_dt(1000)
    a.step(600)
    D = a.dt_history(1000)
    b = sw2d.Sim(z, h, 5.0, dt=1.0, manning=0.05, n_cell=ncell, edge_open=sw2d.EDGE_SOUTH,
                 forcing_mode=sw2d.FORCING_STEP, precision=sw2d.FP32)
    b.set_rain(table, 300.0)
    b.set_dt_schedule(D)
    b.step(600)
    ha, hb = a.read(sw2d.F_H), b.read(sw2d.F_H)
    err = float(np.abs(ha - hb).max())
    sa, sb = a.stats(), b.stats()
    rel = abs(sa["volume"] - sb["volume"]) / sa["volume"]
    say("fp32 max|dh| = %.3e  rel volume = %.3e  t = %.3f/%.3f" % (err, rel, sa["t"], sb["t"]))
    assert err <= 1e-4
    assert rel <= 1e-6


def test_gpu_run_db_dam_break(tmp_path, dam_case):
    """The drop-in driver on a copy of db/dam_break: Results_0.vtk equals the
    reference's file; Results_1.vtk (1500 iterations, t = 60 s) matches the
    oracle at the writer's 6 decimals."""
    say("case: run_db dam_break")
    db = tmp_path / "db"
    shutil.copytree(DAM_BREAK, db / "dam_break")
    (db / "dam_break" / "output").mkdir(exist_ok=True)
    for f in (db / "dam_break" / "output").iterdir():
        f.unlink()
    setup = db / "dam_break" / "input" / "setup.dat"
    setup.write_bytes(setup.read_bytes().replace(b"tday_thour_tmin_tsec: 0 0 100 0",
                                                 b"tday_thour_tmin_tsec: 0 0 2 0"))
    (db / "dir.txt").write_text("dir_number: 1\ndam_break\n")
    sw2d.run_db(str(db))
    out = db / "dam_break" / "output"
    ref = gzip.open(os.path.join(DAM_BREAK, "output", "Results_0.vtk.gz")).read()
    assert (out / "Results_0.vtk").read_bytes() == ref
    assert (out / "WaterLevel.txt").read_text() == "Outlet off\n"
    assert (out / "Results_1.vtk").exists()
    c = dam_case
    d, s = c["dem"], c["setup"]
    o = O.OracleSim(d["z"], (d["z"] != d["nodata"]).astype(np.uint8), c["ic"], d["res"], s["dt"],
                    s["manning_coef"])
    o.set_rain(c["rain"], c["dtrain"])
    o.step(750)
    txt = caseio.vtk_text(c, o.read(O.H), o.read(O.VV), o.read(O.UU), d["z"])
    got = (out / "Results_1.vtk").read_bytes()
    ga, gb = got.split(b"\n"), txt.split(b"\n")
    assert len(ga) == len(gb)
    assert ga[:10] == gb[:10]                      # header and first block label
    num = lambda lines: np.array([float(x) for x in lines if x and not x[:1].isalpha()])
    a, b = num(ga[10:]), num(gb[10:])
    assert a.shape == b.shape
    text_diff = int(np.count_nonzero([x != y for x, y in zip(ga, gb)]))
    say("Results_1: %d lines differ as text (sign of values below 5e-7 or last-digit "
        "rounding), max numeric difference %.2e" % (text_diff, float(np.abs(a - b).max())))
    assert np.abs(a - b).max() <= 1.5e-6
