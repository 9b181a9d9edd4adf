"""Shared fixtures: import paths for the oracle (test infrastructure) and the
ctypes binding of the product library, and the `gpu` marker."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "sw2d-gpu_amd"))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DAM_BREAK = os.path.join(GOLDEN, "dam_break")
PERI_LAKE = os.path.join(GOLDEN, "peri_lake_model")


def materialize_case(src, db_root, name, setup_edits=None):
    """Copies a golden case into db_root/<name>/input (gunzipping *.gz),
    creates its empty output/ directory, applies {key: value} edits to
    setup.dat and returns the case directory."""
    import gzip
    import shutil
    case = os.path.join(db_root, name)
    inp = os.path.join(case, "input")
    os.makedirs(inp, exist_ok=True)
    os.makedirs(os.path.join(case, "output"), exist_ok=True)
    for f in os.listdir(os.path.join(src, "input")):
        a = os.path.join(src, "input", f)
        if f.endswith(".gz"):
            with gzip.open(a, "rb") as g, open(os.path.join(inp, f[:-3]), "wb") as o:
                shutil.copyfileobj(g, o)
        else:
            shutil.copyfile(a, os.path.join(inp, f))
    if setup_edits:
        path = os.path.join(inp, "setup.dat")
        lines = open(path, "rb").read().decode().split("\n")
        for k, v in setup_edits.items():
            for i, ln in enumerate(lines):
                if ln.startswith(k + ":"):
                    lines[i] = "%s: %s%s" % (k, v, "\r" if ln.endswith("\r") else "")
        open(path, "wb").write("\n".join(lines).encode())
    return case


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def dam_case():
    import caseio
    return caseio.read_case(DAM_BREAK)


def dam_break_1d(rows=100, cols=200, transpose=False):
    """Dam break on a flat bed: h = 1 for x < 100 m, 0.1 beyond (dx = 1)."""
    z = np.zeros((rows, cols))
    h = np.where(np.arange(cols)[None, :] < 100, 1.0, 0.1) * np.ones((rows, cols))
    if transpose:
        return z.T.copy(), h.T.copy()
    return z, h


def random_case(seed, rows, cols, holes=0.05, relief=2.0, wet=0.7):
    rng = np.random.default_rng(seed)
    z = rng.uniform(0.0, relief, (rows, cols))
    h = rng.uniform(0.0, 1.5, (rows, cols)) * (rng.uniform(size=(rows, cols)) < wet)
    mask = rng.uniform(size=(rows, cols)) < holes
    z[mask] = -9999.0
    h[mask] = 0.0
    return z, h


def ensure_oracle():
    import oracle_py
    if not os.path.exists(oracle_py.LIB):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
