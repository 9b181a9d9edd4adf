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
    h300, vv300, uu300 = o.read(O.H), o.read(O.VV), o.read(O.UU)
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
    # every output file byte for byte as the oracle's state, written by the
    # restated writer: the FP64 state is bitwise the oracle's
    for k, (hh, v, u) in enumerate([(ref["ic"], None, None), (h150, vv150, uu150), (h300, vv300, uu300)]):
        if k == 0:
            continue  # Results_0: the first iteration's velocities; covered by the dam_break test
        want = caseio.vtk_text(ref, hh, v, u, ref["dem"]["z"])
        got = open(os.path.join(out, "Results_%d.vtk" % k), "rb").read()
        assert got == want, "Results_%d.vtk" % k
    # WaterLevel.txt: the initial gauge depths, then one line per output
    cells = sorted(caseio.gauge_cells(ref))  # printed in raster order
    want = "".join("".join(" %f" % x for x in hh.ravel()[cells]) + "\n"
                   for hh in (ref["ic"], h150, h300))
    assert open(os.path.join(out, "WaterLevel.txt")).read() == want
