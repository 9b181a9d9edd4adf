"""The seeded workload generators reproduce the committed checksums
(tests/golden/generator_checksums.json, written by
tests/golden/generator_fixtures.py): configurations 2-5 exist only through
these generators, so any change to them would silently change every
benchmark and GPU test."""
import json
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import generator_fixtures as GF  # noqa: E402


def test_generators_match_committed_checksums():
    with open(GF.PATH) as f:
        want = json.load(f)
    got = GF.compute()
    assert sorted(got) == sorted(want)
    bad = [k for k in want if got[k]["sha256_16"] != want[k]["sha256_16"] or got[k]["shape"] != want[k]["shape"]]
    assert not bad, bad[:10]


def test_fixture_statistics_are_plausible():
    """The recorded shapes and ranges are those of the configurations."""
    with open(GF.PATH) as f:
        want = json.load(f)
    assert want["config2_z"]["shape"] == [1024, 1024] and want["config3_z"]["shape"] == [8192, 8192]
    assert 0.03 <= want["config3_n"]["min"] and want["config3_n"]["max"] <= 0.10
    assert want["config3_z"]["max"] <= 100.0 + 1e-9 and want["config3_z"]["min"] >= -1e-9
    assert want["config4_sides"]["min"] >= 256 and want["config4_sides"]["max"] <= 2048
    assert want["config5_sides"]["min"] >= 128 and want["config5_sides"]["max"] <= 4096
    assert 600 <= want["config5_durations"]["min"] and want["config5_durations"]["max"] <= 21600
