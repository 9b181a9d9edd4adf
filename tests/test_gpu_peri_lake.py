"""The reference's real-catchment case db/peri_lake_model (561 x 449 cells
at 12.5 m, about half NODATA, evaporation on, one Q sink, four gauges,
dt = 0.04 s) through the product: the case reader, the device forcing
(rain, point sink, evaporation table) and the drop-in driver.  The case's
own end time is the whole rain table (~34 days, 3.7e7 iterations); the
test runs 5 minutes (3750 iterations) from 05:01, inside an hour with rain
and daylight (so rain, evaporation and the sink all act) and off the table
breakpoints, where the reference's forcing formula divides by zero.
The reference's outputs of this case are missing from its repository, so
the check is the FP64 oracle on the same inputs."""
import os
import sys

import numpy as np
import pytest

from conftest import PERI_LAKE, materialize_case

pytestmark = pytest.mark.gpu

import caseio  # noqa: E402
import oracle_py as O  # noqa: E402
import sw2d  # noqa: E402

TOL64 = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_for(ref):
    s, dem = ref["setup"], ref["dem"]
    act = (dem["z"] != dem["nodata"]).astype(np.uint8)
    o = O.OracleSim(dem["z"], act, ref["ic"], dem["res"], s["dt"], s["manning_coef"], INT=s["INT"],
                    INF=s["INF"], LWL=s["LWL"], EV_WL_min=s["EV_WL_min"],
                    evaporation_on=s["evaporation_on"], t0=ref["time0"])
    o.set_rain(ref["rain"], ref["dtrain"])
    o.set_sources(ref["q"], s["dtoq"], caseio.source_cells(ref))
    o.set_evaporation([O.evap_entry(ref["albedo"], t, r, ref["dtrain"])
                       for t, r in zip(ref["T"], ref["Rg"])])
    return o


def field_diffs(o, g):
    return {k: float(np.abs(o.read(a) - g.read(b)).max())
            for k, a, b in [("h", O.H, sw2d.F_H), ("U", O.U, sw2d.F_U), ("V", O.V, sw2d.F_V)]}


def test_gpu_peri_lake_fp64_and_drop_in(tmp_path):
    say("case: peri_lake_model 5 min FP64 (3750 iterations) + drop-in driver")
    db = tmp_path / "db"
    d = materialize_case(PERI_LAKE, str(db), "peri_lake_model",
                         {"tday0_thour0_tmin0_tsec0": "0 5 1 0", "tday_thour_tmin_tsec": "0 5 6 0",
                          "dkout": "150", "dpout": "150"})
    (db / "dir.txt").write_text("dir_number: 1\nperi_lake_model\n")
    ref = caseio.read_case(d)
    assert ref["timmax"] - ref["time0"] == 300.0
    o = oracle_for(ref)
    g = sw2d.Case(d).create()

    o.step(1875)
    g.step(1875)
    dd = field_diffs(o, g)
    say("t = 150 s: %s" % dd)
    assert max(dd.values()) <= TOL64, dd
    h150, vv150, uu150 = o.read(O.H), o.read(O.VV), o.read(O.UU)
    st = o.stats()
    say("stats at 150 s: %s" % st)
    assert st["evap_out"] > 0 and st["source_in"] < 0 and st["rain_in"] > 0

    o.step(1875)
    g.step(1875)
    dd = field_diffs(o, g)
    say("t = 300 s: %s" % dd)
    assert max(dd.values()) <= TOL64, dd
    sa, sg = o.stats(), g.stats()
    for k in ("volume", "rain_in", "source_in", "evap_out", "clamp_gain"):
        assert abs(sa[k] - sg[k]) <= 1e-9 * max(1.0, abs(sa[k])), (k, sa[k], sg[k])

    # the drop-in driver on the same case directory
    sw2d.run_db(str(db))
    out = os.path.join(d, "output")
    names = sorted(os.listdir(out))
    say("outputs: %s" % names)
    for f in ("Results_0.vtk", "Results_1.vtk", "WaterLevel.txt"):
        assert f in names
    assert any(n.startswith("TimeSimu_") for n in names)
    txt = caseio.vtk_text(ref, h150, vv150, uu150, ref["dem"]["z"])
    got = open(os.path.join(out, "Results_1.vtk"), "rb").read()
    ga, gb = got.split(b"\n"), txt.split(b"\n")
    assert len(ga) == len(gb) and ga[:10] == gb[:10]
    num = lambda lines: np.array([float(x) for x in lines if x and not x[:1].isalpha()])
    a, b = num(ga[10:]), num(gb[10:])
    assert a.shape == b.shape and np.abs(a - b).max() <= 1.5e-6
    # WaterLevel.txt: the initial gauge depths, then one line per output
    wl = [ln.split() for ln in open(os.path.join(out, "WaterLevel.txt")).read().splitlines()]
    assert len(wl) >= 2 and all(len(x) == 4 for x in wl)
    cells = sorted(caseio.gauge_cells(ref))  # printed in raster order
    assert np.allclose([float(v) for v in wl[0]], ref["ic"].ravel()[cells], atol=1.5e-6)
    assert np.allclose([float(v) for v in wl[1]], h150.ravel()[cells], atol=1.5e-6)
