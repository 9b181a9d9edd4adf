"""Seed -> checksum fixtures of the synthetic workload generators.

The benchmark configurations 2-5 exist only through the library's seeded
generators (include/sw2d_gen.h, sw2d-gpu_amd/workloads.py): no reference data
holds them.  This module computes one SHA-256 digest (first 16 hex digits)
per generated array, over its float64 bytes, together with its shape, min,
max and mean.  `python tests/golden/generator_fixtures.py` rewrites
generator_checksums.json; tests/test_generator_fixtures.py recomputes every
entry and compares, so a change of a generator changes a committed file.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "sw2d-gpu_amd"))
import sw2d  # noqa: E402
import workloads  # noqa: E402

PATH = os.path.join(HERE, "generator_checksums.json")


def digest(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return {"sha256_16": hashlib.sha256(a.tobytes()).hexdigest()[:16], "shape": list(a.shape),
            "min": float(a.min()), "max": float(a.max()), "mean": float(a.mean())}


def entries():
    """name -> array, for every generator primitive and configuration input."""
    out = {
        "fractal_dem_65x97_seed7": sw2d.gen_fractal_dem(65, 97, 7, 0.6, 0.0, 50.0),
        "tilt_south_65x97": sw2d.gen_tilt(sw2d.gen_fractal_dem(65, 97, 7, 0.6, 0.0, 50.0), 5.0, 0.005,
                                          sw2d.EDGE_SOUTH),
        "uniform_1000_seed11": sw2d.gen_uniform((1000,), 11, 0.03, 0.10),
        "lognormal_1000_seed12": sw2d.gen_lognormal((1000,), 12, 0.0, 0.5),
        "gamma_1000_seed13": sw2d.gen_gamma((1000,), 13, 0.8, 20.0),
    }
    w = workloads.config1()
    out["config1_z"], out["config1_h0"] = w["z"], w["h0"]
    w = workloads.config2()
    out["config2_z"], out["config2_n"], out["config2_rain"] = w["z"], w["n_cell"], w["rain"][0]
    w = workloads.config3()
    out["config3_z"], out["config3_n"], out["config3_h0"] = w["z"], w["n_cell"], w["h0"]
    out["config3_rain_map"] = w["rain"][2]
    w = workloads.config4()
    out["config4_sides"] = np.array([g["side"] for g in w["grids"]], dtype=np.float64)
    for g, grid in enumerate(w["grids"]):
        out["config4_grid%02d_z" % g] = grid["z"]
        out["config4_grid%02d_n" % g] = grid["n_cell"]
        out["config4_grid%02d_hyetographs" % g] = np.array(grid["hyetographs"])
    w = workloads.config5(n_requests=48)
    r = w["requests"]
    out["config5_arrivals"] = np.array([q["arrival"] for q in r])
    out["config5_sides"] = np.array([q["side"] for q in r], dtype=np.float64)
    out["config5_durations"] = np.array([q["duration"] for q in r])
    for k in range(0, 48, 8):
        out["config5_req%02d_z" % k] = r[k]["z"]
        out["config5_req%02d_n" % k] = r[k]["n_cell"]
        out["config5_req%02d_rain" % k] = r[k]["rain"]
    return out


def compute():
    return {k: digest(v) for k, v in entries().items()}


if __name__ == "__main__":
    with open(PATH, "w") as f:
        json.dump(compute(), f, indent=1, sort_keys=True)
    print("wrote", PATH)
