"""Synthetic benchmark workloads (host generators only, no GPU): seeds
reproduce, and the drawn sizes, durations and arrival times stay inside the
ranges the ensemble and serving configurations define."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "sw2d-gpu_amd"))
import workloads  # noqa: E402


def test_ensemble_grids_reproduce_and_stay_in_range():
    a, b = workloads.config4(n_grids=3), workloads.config4(n_grids=3)
    for ga, gb in zip(a["grids"], b["grids"]):
        assert 256 <= ga["side"] <= 2048 and ga["z"].shape == (ga["side"], ga["side"])
        assert np.array_equal(ga["z"], gb["z"]) and np.array_equal(ga["n_cell"], gb["n_cell"])
        assert len(ga["hyetographs"]) == 4
        for ha, hb in zip(ga["hyetographs"], gb["hyetographs"]):
            assert ha.shape == (24,) and np.array_equal(ha, hb) and (ha >= 0).all()
        assert 0.03 <= ga["n_cell"].min() and ga["n_cell"].max() <= 0.10
    assert a["params"]["t_end"] == 7200.0


def test_serving_stream_reproduces_and_stays_in_range():
    a, b = workloads.config5(n_requests=6), workloads.config5(n_requests=6)
    arr = [r["arrival"] for r in a["requests"]]
    assert arr == sorted(arr) and arr[0] > 0
    assert arr == [r["arrival"] for r in b["requests"]]
    for r in a["requests"]:
        assert 128 <= r["side"] <= 4096
        assert 600.0 <= r["duration"] <= 21600.0
        assert len(r["rain"]) * 300.0 >= r["duration"]
        assert 0.02 <= r["n_cell"].min() and r["n_cell"].max() <= 0.12
