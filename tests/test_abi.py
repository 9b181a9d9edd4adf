"""The C ABI library loads on a machine without a GPU, exports every entry
point include/sw2d.h declares, and its host-only parts (case reader, VTK
writer, forcing formulas) agree with the oracle's restatements."""
import gzip
import os
import re
import shutil

import numpy as np
import pytest

from conftest import materialize_case, PERI_LAKE, DAM_BREAK, ROOT, ensure_oracle

ensure_oracle()
import caseio  # noqa: E402
import oracle_py as O  # noqa: E402
import sw2d  # noqa: E402


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(sw2d_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = sw2d.lib()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(sw2d.EXPORTS)


def test_stream_probe_library_loads():
    """The streaming-ceiling probe (bench.py --stream-ceiling) is built beside
    the library and exports its entry point; with a bad argument it refuses
    before touching the device."""
    import ctypes as C
    path = os.path.join(ROOT, "sw2d-gpu_amd", "libsw2d_probe.so")
    L = C.CDLL(path)
    f = L.sw2d_probe_stream5r3w
    f.argtypes = [C.POINTER(C.c_void_p), C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_int,
                  C.POINTER(C.c_double)]
    ms = C.c_double()
    assert f(None, 1024, 1, 64, 0, 1, C.byref(ms)) == -1


def test_case_reader_matches_restatement(dam_case):
    c = sw2d.Case(DAM_BREAK)
    I = c.info
    ref = dam_case
    assert (I.rows, I.cols) == ref["dem"]["z"].shape
    assert I.cellsize == ref["dem"]["res"] and I.nodata == ref["dem"]["nodata"]
    assert I.dt == ref["setup"]["dt"] and I.manning_coef == ref["setup"]["manning_coef"]
    assert I.timmax == ref["timmax"] == 6000.0
    assert np.array_equal(c.array(sw2d.CASE_DEM), ref["dem"]["z"].ravel())
    assert np.array_equal(c.array(sw2d.CASE_IC), ref["ic"].ravel())
    assert np.array_equal(c.array(sw2d.CASE_RAIN), ref["rain"])
    assert np.array_equal(c.array(sw2d.CASE_Q), ref["q"].ravel())
    assert I.n_q_rows == 817 and I.n_rain == 120 and I.dtrain == 60.0
    assert list(c.source_cells()) == caseio.source_cells(ref) == [10001]


def test_case_reader_peri_lake(tmp_path):
    """The reference's real-catchment case (NODATA mask, evaporation,
    a Q sink, four gauges, end time from the rain table): the product reader
    agrees with the restatement field by field."""
    d = materialize_case(PERI_LAKE, str(tmp_path), "peri_lake_model")
    c, ref = sw2d.Case(d), caseio.read_case(d)
    I = c.info
    assert (I.rows, I.cols) == ref["dem"]["z"].shape == (561, 449)
    assert I.cellsize == 12.5 and I.dt == 0.04 and I.evaporation_on == 1
    assert I.timmax == ref["timmax"] == (ref["cont_rain"] - 2) * 3600.0
    assert np.array_equal(c.array(sw2d.CASE_DEM), ref["dem"]["z"].ravel())
    assert np.array_equal(c.array(sw2d.CASE_IC), ref["ic"].ravel())
    assert np.array_equal(c.array(sw2d.CASE_RAIN), ref["rain"])
    assert np.array_equal(c.array(sw2d.CASE_Q), ref["q"].ravel())
    assert list(c.source_cells()) == caseio.source_cells(ref)
    ev = [O.evap_entry(ref["albedo"], t, r, ref["dtrain"]) for t, r in zip(ref["T"], ref["Rg"])]
    assert np.array_equal(c.array(sw2d.CASE_EVAPO), ev)
    act = (ref["dem"]["z"] != ref["dem"]["nodata"]).mean()
    assert 0.3 < act < 0.7  # about half of the raster is NODATA


def test_vtk_writer_bytes_match_reference_output(dam_case, tmp_path):
    """The product writer, fed the oracle's first-iteration velocities,
    reproduces the reference's committed Results_0.vtk byte for byte."""
    c = dam_case
    d, s = c["dem"], c["setup"]
    act = (d["z"] != d["nodata"]).astype(np.uint8)
    o = O.OracleSim(d["z"], act, c["ic"], d["res"], s["dt"], s["manning_coef"])
    o.step(1)
    case = sw2d.Case(DAM_BREAK)
    out = str(tmp_path / "Results_0.vtk")
    case.write_vtk(out, c["ic"], o.read(O.VV), o.read(O.UU), d["z"])
    ref = gzip.open(os.path.join(DAM_BREAK, "output", "Results_0.vtk.gz")).read()
    assert open(out, "rb").read() == ref


def test_forcing_formulas_match_oracle():
    a = np.array([2.0, 4.0, 8.0, 8.0, 1.0])
    for t in [0.0, 0.3, 1.0, 1.5, 59.9, 60.0, 61.0, 125.5, 299.0, 1e6]:
        for mode in (0, 1):
            x, y = sw2d.forcing_value(t, a, 60.0, mode), O.stream_rate(t, a, 60.0, mode)
            assert (x == y) or (np.isnan(x) and np.isnan(y))
    T = np.linspace(5, 35, 31)
    Rg = np.linspace(-50, 900, 31)
    e = sw2d.evaporation_table(0.25, T, Rg, 3600.0)
    assert np.array_equal(e, [O.evap_entry(0.25, t, r, 3600.0) for t, r in zip(T, Rg)])


def test_missing_file_is_an_io_error(tmp_path):
    d = tmp_path / "case"
    shutil.copytree(DAM_BREAK, d)
    os.remove(d / "input" / "rain.dat")
    with pytest.raises(sw2d.SW2DError) as e:
        sw2d.Case(str(d))
    assert e.value.code == sw2d.E_IO and "rain.dat" in str(e.value)


def test_malformed_setup_is_rejected(tmp_path):
    d = tmp_path / "case"
    shutil.copytree(DAM_BREAK, d)
    p = d / "input" / "setup.dat"
    p.write_bytes(p.read_bytes().replace(b"dpout:", b"dpoutx:"))
    with pytest.raises(sw2d.SW2DError) as e:
        sw2d.Case(str(d))
    assert e.value.code == sw2d.E_IO


def test_footprint_is_about_32_bytes_per_cell_fp32():
    # h, U, V ping-pong pairs (24 B), bed (4 B) and Manning raster (4 B), the
    # active-cell mask (1 B) and a staging buffer of at most 32 MiB
    b = sw2d.footprint_bytes(8192, 8192, sw2d.FP32, True)
    per = b / (8192 * 8192)
    assert 33 <= per <= 34.5
    assert b - sw2d.footprint_bytes(8192, 8192, sw2d.FP32, False) >= 4 * 8192 * 8192
    # the staging buffer is bounded: doubling the rows adds the arrays only
    b2 = sw2d.footprint_bytes(16384, 8192, sw2d.FP32, True)
    assert b2 - b < 34 * 8192 * 8192


def test_library_friction_pow_is_the_oracle_sequence():
    """sw2d_friction_pow (the sequence the FP64 step evaluates on the device)
    gives the oracle's bits on every argument."""
    import random
    import oracle_py as O
    rnd = random.Random(9)
    for _ in range(20000):
        x = 10 ** rnd.uniform(-4, 3)
        assert sw2d.friction_pow(x) == O.pow_fric(x), x


def test_fp64_form_refuses_g_th_below_its_square_root_range():
    """The FP64 step takes square roots of g h for h > th with a sequence
    valid from 2^-767 on: sw2d_create refuses g th < 2^-760 (before it looks
    for a device), and accepts the reference's th = 1e-4 down to g th =
    2^-760 as far as the argument checks go."""
    z = np.zeros((8, 8))
    for th in (0.0, 1e-300):
        with pytest.raises(sw2d.SW2DError) as e:
            sw2d.Sim(z, z, 1.0, precision=sw2d.FP64, th=th)
        assert e.value.code == sw2d.E_INVALID and "th" in str(e.value)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except ImportError:
        has_gpu = False
    if not has_gpu:
        for th in (1e-4, 2.0 ** -760 / 9.80 * 1.0000001):
            with pytest.raises(sw2d.SW2DError) as e:
                sw2d.Sim(z, z, 1.0, precision=sw2d.FP64, th=th)
            assert e.value.code == sw2d.E_DEVICE


def test_fp32_form_refuses_a_threshold_below_a_normal_float():
    """sw2d_create checks its arguments before it looks for a device: the
    FP32 form refuses th < 1e-30 (its friction power needs face depths above
    th to be normal floats); with th = 1e-4 the same call passes the checks
    and, on a machine without a GPU, fails only for the missing device."""
    z = np.zeros((8, 8))
    with pytest.raises(sw2d.SW2DError) as e:
        sw2d.Sim(z, z, 1.0, precision=sw2d.FP32, th=0.0)
    assert e.value.code == sw2d.E_INVALID and "th" in str(e.value)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except ImportError:
        has_gpu = False
    if not has_gpu:
        with pytest.raises(sw2d.SW2DError) as e:
            sw2d.Sim(z, z, 1.0, precision=sw2d.FP32, th=1e-4)
        assert e.value.code == sw2d.E_DEVICE
