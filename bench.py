"""Benchmark of the SW2D time step on one MI355X.

Workload (default): the large single grid -- 8192^2 cells at dx = 2 m,
seeded fractal bed 0-100 m, 10 m reservoir of 512^2 cells released at t = 0,
30 mm/h x lognormal(0, 0.5) rain per 64^2 block, Manning n ~ U(0.03, 0.10)
per cell, CFL 0.5; FP32 state with FP64 volume accumulators.  A step is one
iteration of the reference loop (one update of the whole grid).

Prints one JSON line: Gcell-updates/s over the timed steps (state resident
in HBM), the HBM roofline of the step kernels measured with HIP events on the
library's stream, and a CPU baseline (the FP64 OpenMP oracle on the same
state, a bounded number of iterations).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sw2d-gpu_amd"))

METRIC = "Gcell-updates/s & % HBM roofline; ensemble scenarios/s + p99 latency"
HBM_PEAK_GBS = 8000.0
# compulsory HBM bytes per cell-update (read h, U, V, z, n; write h, U, V)
B_ALG = {"fp32": 32.0, "fp64": 64.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_per_launch(size, precision):
    """HBM bytes per k_step launch on this workload, measured with two
    rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by what the same
    counters read for a plain copy kernel, as recorded in
    profiles/traffic.json; None when no measurement is recorded."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f).get("%s_%d" % (precision, size))
    except (OSError, ValueError):
        return None


def cpu_baseline(w, h, U, V, seconds=15.0, max_iters=40):
    """FP64 OpenMP oracle on the same grid and state (test infrastructure,
    timed as the CPU baseline)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    p = w["params"]
    act = np.ones(w["z"].shape, np.uint8)
    o = O.OracleSim(w["z"], act, h, w["dx"], p["dt"], p["manning"], cfl=p.get("cfl", 0.0),
                    edge_open=p.get("edge_open", 0), forcing_mode=p.get("forcing_mode", 0),
                    n_cell=w["n_cell"])
    table, dtrain, bmap, block = w["rain"]
    o.set_rain(table, dtrain, bmap, block)
    o.set_state(h, U, V)
    o.step(1)  # first touch of the oracle's arrays, untimed
    n, t0 = 0, time.perf_counter()
    while n < max_iters:
        o.step(1)
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    cells = w["z"].size
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=cells * n / dt / 1e9, unit="Gcell-updates/s", cores=threads, kind="port",
                sample="%d FP64 iterations of the same %dx%d grid from the GPU's warmed-up state "
                       "(%.1f s)" % (n, w["z"].shape[0], w["z"].shape[1], dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is a design "
              "(row strips with a 3-row halo), not implemented", file=sys.stderr)
        sys.exit(2)

    import sw2d
    import workloads

    prec = sw2d.FP32 if args.precision == "fp32" else sw2d.FP64
    t0 = time.perf_counter()
    w = workloads.config3(size=args.size, precision=prec)
    log("workload built in %.1f s: %s" % (time.perf_counter() - t0, w["desc"]))
    s = workloads.make_sim(w)
    cells = w["z"].size

    s.step(args.warmup)
    s.sync()
    st0 = s.stats()
    log("warmup %d iterations done: t = %.2f s, D = %.4f s" % (args.warmup, st0["t"], st0["last_D"]))

    # timed region: the library stream is synchronised on both sides; the
    # iterations run as replayed HIP graphs of the two launches
    s.sync()
    t1 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    wall = time.perf_counter() - t1
    st1 = s.stats()
    log("timed %d iterations in %.3f s: t = %.2f s" % (args.steps, wall, st1["t"]))
    value = cells * args.steps / wall / 1e9

    # kernel time of the stencil step: HIP events on the library stream
    # around each launch of a further sample of iterations
    nprof = max(20, min(args.steps, 200))
    s.profile_begin(nprof)
    s.step(nprof)
    s.sync()
    ms, nprof = s.profile_end()
    kern_ms = ms[1] / max(nprof, 1)
    achieved = B_ALG[args.precision] * cells / (kern_ms * 1e-3) / 1e9

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if prec == sw2d.FP32 else "f64",
        "data": "synthetic (seeded fractal DEM, reservoir release, lognormal block rain)",
        "config": {"workload": "large single grid %dx%d dx=2 m, FP32 state + FP64 accumulators, "
                               "CFL 0.5" % (args.size, args.size),
                   "cells": cells, "precision": args.precision},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_per_launch(args.size, args.precision),
                     "kernel": "k_step (one iteration)", "bytes_per_cell": B_ALG[args.precision],
                     "ms_step": round(kern_ms, 4), "ms_control": round(ms[0] / max(nprof, 1), 4),
                     "profiled_iterations": nprof},
        "sim_time_s": [round(st0["t"], 3), round(st1["t"], 3)],
    }

    h = U = V = None
    if not args.no_cpu or not args.no_fp64:
        h, U, V = s.read(sw2d.F_H), s.read(sw2d.F_U), s.read(sw2d.F_V)
    s.close()
    if not args.no_fp64 and prec == sw2d.FP32:
        s64 = workloads.make_sim(w, precision=sw2d.FP64)
        s64.write_state(h, U, V)
        n64 = max(50, args.steps // 10)
        s64.step(5)
        s64.sync()
        t2 = time.perf_counter()
        s64.step(n64)
        s64.sync()
        w64 = time.perf_counter() - t2
        s64.profile_begin(50)
        s64.step(50)
        ms64, np64 = s64.profile_end()
        s64.close()
        out["fp64_value"] = round(cells * n64 / w64 / 1e9, 3)
        out["fp64_roofline_frac"] = round(
            B_ALG["fp64"] * cells / (ms64[1] / np64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        log("fp64: %.3f Gcell-updates/s" % out["fp64_value"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, h, U, V)
        log("cpu baseline: %s" % out["cpu_baseline"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
