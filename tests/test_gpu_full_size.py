"""Full-size checks on the 8192 x 8192 workload the benchmark times: dx = 2 m,
fractal relief 0-100 m, Manning raster, a 10 m reservoir of 512^2 cells
released at t = 0, 30 mm/h rain times a lognormal multiplier per 64^2 block,
CFL 0.5.

* FP64 against the FP64 oracle (OpenMP) over 8 iterations, from the
  workload's state plus a 2 cm film on every cell, so every face of the
  grid carries flow over the rough bed: max |dh|, |dU|, |dV| <= 1e-10 and the
  same rain, outflow and volume sums.
* FP32 over 300 iterations from the workload's own state, through
  properties that do not depend on the size: every value finite, the volume
  closes against the FP64 accounting sums, and the state is bitwise the same
  for two segment heights.
"""
import sys

import numpy as np
import pytest

from conftest import ensure_oracle

pytestmark = pytest.mark.gpu

ensure_oracle()
import oracle_py as O  # noqa: E402
import sw2d  # noqa: E402
import workloads  # noqa: E402

N = 8192
TOL64 = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gpu_full_size_fp64_matches_oracle():
    say("case: 8192^2 workload FP64 with a 2 cm film, 8 iterations against the oracle")
    w = workloads.config3(N, precision=sw2d.FP64)
    z, n, dx = w["z"], w["n_cell"], w["dx"]
    h0 = w["h0"] + 0.02
    p = w["params"]
    table, dtrain, bmap, block = w["rain"]
    act = (z != -9999.0).astype(np.uint8)
    o = O.OracleSim(z, act, h0, dx, p["dt"], p["manning"], n_cell=n, cfl=p["cfl"],
                    forcing_mode=p["forcing_mode"])
    g = sw2d.Sim(z, h0, dx, n_cell=n, **p)
    o.set_rain(table, dtrain, bmap, block)
    g.set_rain(table, dtrain, bmap, block)
    o.step(8)
    g.step(8)
    g.sync()
    for name, fo, fg in (("h", O.H, sw2d.F_H), ("U", O.U, sw2d.F_U), ("V", O.V, sw2d.F_V)):
        a, b = o.read(fo), g.read(fg)
        d = float(np.max(np.abs(a - b)))
        say("%s max diff %.3e, %d cells differ" % (name, d, int(np.count_nonzero(a != b))))
        assert d <= TOL64, (name, d)
        del a, b
    so, sg = o.stats(), g.stats()
    say("t %.6f / %.6f" % (so["t"], sg["t"]))
    assert so["t"] == sg["t"]
    for k in ("volume", "rain_in", "outflow", "clamp_gain"):
        assert abs(so[k] - sg[k]) <= 1e-9 * max(1.0, abs(so[k])), (k, so[k], sg[k])
    g.close()
    del o


def test_gpu_full_size_fp32_properties(monkeypatch):
    say("case: 8192^2 workload FP32, 300 iterations, default and 40-row segments")
    w = workloads.config3(N, precision=sw2d.FP32)
    table, dtrain, bmap, block = w["rain"]
    v0 = float(np.sum(w["h0"])) * w["dx"] * w["dx"]
    runs = []
    for rows in (None, 40):
        if rows is None:
            monkeypatch.delenv("SW2D_SEG_ROWS", raising=False)
        else:
            monkeypatch.setenv("SW2D_SEG_ROWS", str(rows))
        g = sw2d.Sim(w["z"], w["h0"], w["dx"], n_cell=w["n_cell"], **w["params"])
        g.set_rain(table, dtrain, bmap, block)
        g.step(300)
        g.sync()
        st = g.stats()
        fields = [g.read(f) for f in (sw2d.F_H, sw2d.F_U, sw2d.F_V)]
        g.close()
        for f in fields:
            assert np.all(np.isfinite(f))
        assert np.all(fields[0] >= 0.0)  # no inactive cells in this workload
        inflow = st["rain_in"] + st["source_in"] - st["evap_out"] - st["outflow"] + st["clamp_gain"]
        closure = abs(st["volume"] - (v0 + inflow)) / st["volume"]
        say("segments %s: t %.4f s, volume %.6e m3, rain %.6e m3, closure %.2e"
            % (rows or "default", st["t"], st["volume"], st["rain_in"], closure))
        assert st["iterations"] == 300
        assert closure <= 1e-6
        runs.append((fields, st["t"]))
    monkeypatch.delenv("SW2D_SEG_ROWS", raising=False)
    (fa, ta), (fb, tb) = runs
    assert ta == tb
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y), float(np.abs(x - y).max())
