"""Pins of the FP64 oracle (oracle/sw2d_oracle.c) against the reference's
golden data, an independent literal transcription and closed forms.

Reference: SW2D-GPU_linux/src/SW2D-GPU.cu (L:n below)."""
import gzip
import math
import os

import numpy as np
import pytest

from conftest import DAM_BREAK, dam_break_1d, ensure_oracle, random_case

ensure_oracle()
import caseio  # noqa: E402
import oracle_py as O  # noqa: E402
import sw2d_np as NP  # noqa: E402

G, TH = 9.80, 1.0e-4


def run_pair(z, h0, dx, dt, n, iters, rain=None, **kw):
    act = (z != -9999.0).astype(np.uint8)
    o = O.OracleSim(z, act, h0, dx, dt, n, **kw)
    r = NP.RefLayoutSim(z, h0, -9999.0, dx, dt, n, **{k: v for k, v in kw.items()
                                                     if k in ("INT", "INF", "LWL", "EV_WL_min")})
    if rain is not None:
        table, dtrain = rain
        o.set_rain(table, dtrain)
    for _ in range(iters):
        o.step(1)
        if rain is not None:
            r.iterate(lambda t: O.stream_rate(t, table, dtrain, 0) / (dtrain * 1000.0))
        else:
            r.iterate()
    return o, r


def assert_same(o, r):
    pairs = [(o.read(O.H), r.cells(r.h)), (o.read(O.U), r.u_faces()), (o.read(O.V), r.v_faces()),
             (o.read(O.UU), r.cells(r.uu)), (o.read(O.VV), r.cells(r.vv))]
    for a, b in pairs:
        assert a.shape == b.shape
        assert np.array_equal(a, b), np.abs(a - b).max()


def test_kat_results0_vtk(dam_case):
    """After one iteration on db/dam_break the reference wrote
    output/Results_0.vtk (L:1339-1472): depth = initial condition, the
    breach faces carry rule-2 velocities.  Byte-for-byte equality."""
    c = dam_case
    d, s = c["dem"], c["setup"]
    act = (d["z"] != d["nodata"]).astype(np.uint8)
    o = O.OracleSim(d["z"], act, c["ic"], d["res"], s["dt"], s["manning_coef"], INT=s["INT"],
                    INF=s["INF"], LWL=s["LWL"], EV_WL_min=s["EV_WL_min"])
    o.set_rain(c["rain"], c["dtrain"])
    o.set_sources(c["q"], s["dtoq"], caseio.source_cells(c))
    o.step(1)
    txt = caseio.vtk_text(c, c["ic"], o.read(O.VV), o.read(O.UU), d["z"])
    ref = gzip.open(os.path.join(DAM_BREAK, "output", "Results_0.vtk.gz")).read()
    assert txt == ref
    # the 22 nonzero velocities are the rule-2 closed form, halved by uu_vv
    uu = o.read(O.UU)
    expect = 0.35 * 5 * math.sqrt(2 * G * 5) / 2.5 / 2
    assert np.count_nonzero(uu) == 22
    assert np.allclose(np.abs(uu[uu != 0]), expect, rtol=1e-15)


@pytest.mark.parametrize("seed,shape", [(1, (13, 17)), (2, (9, 31)), (3, (24, 8)), (4, (5, 5))])
def test_oracle_matches_literal_transcription(seed, shape):
    """Two independent restatements (explicit faces with walls vs the
    reference's flat N = R*C + C/2 layout with aliasing) agree bitwise."""
    z, h = random_case(seed, *shape)
    o, r = run_pair(z, h, 1.0, 0.01, 0.03, 150)
    assert_same(o, r)


def test_oracle_matches_literal_on_dam_break(dam_case):
    c = dam_case
    d, s = c["dem"], c["setup"]
    o, r = run_pair(d["z"], c["ic"], d["res"], s["dt"], s["manning_coef"], 120)
    assert_same(o, r)


def test_oracle_matches_literal_with_rain():
    z, h = random_case(7, 11, 14)
    table = np.array([3.0, 5.0, 5.0, 1.0, 0.0, 2.0])
    o, r = run_pair(z, h, 2.0, 0.05, 0.05, 200, rain=(table, 7.0), INT=0.45, INF=0.2,
                    LWL=0.3, EV_WL_min=0.05)
    assert_same(o, r)


def test_rule2_face_flux_closed_form(dam_case):
    """Rule 2 (L:149-155) at the breach: U = 0.35*hh3*sqrt(2 g hh3)."""
    c = dam_case
    d, s = c["dem"], c["setup"]
    act = (d["z"] != d["nodata"]).astype(np.uint8)
    o = O.OracleSim(d["z"], act, c["ic"], d["res"], s["dt"], s["manning_coef"])
    o.step(1)
    U = o.read(O.U)
    val = 0.350 * 5.0 * math.sqrt(2.00 * G * 5.0)
    assert abs(val - 17.3241161) < 1e-7          # stale Windows fixture value
    # faces between DEM rows 13 and 14 at the breach columns 44..54
    assert np.all(U[14, 44:55] == val)
    assert np.count_nonzero(U) == 11


def test_rule1_closed_form():
    """Rule 1a/1b (L:130-145): a wet cell above a step discharges
    0.544 h sqrt(g h) over the edge."""
    z = np.array([[5.0], [0.0]])
    h = np.array([[2.0], [0.0]])
    o = O.OracleSim(z, np.ones_like(z, dtype=np.uint8), h, 1.0, 0.001, 0.0)
    o.step(1)
    U = o.read(O.U)
    assert U[1, 0] == 0.5440 * 2.0 * math.sqrt(G * 2.0)
    # mirror: wet cell below the step, flow goes north (negative)
    o2 = O.OracleSim(z[::-1].copy(), np.ones((2, 1), dtype=np.uint8), h[::-1].copy(), 1.0, 0.001, 0.0)
    o2.step(1)
    assert o2.read(O.U)[1, 0] == -0.544 * 2.0 * math.sqrt(G * 2.0)


def test_lake_at_rest_is_exact():
    """Flat free surface over an uneven bed, all wet: the gravity term is
    exactly zero and every face stays 0 bitwise."""
    rng = np.random.default_rng(5)
    z = rng.integers(0, 1024, (20, 30)) / 1024.0   # dyadic: h + z == 2 exactly
    h = 2.0 - z
    o = O.OracleSim(z, np.ones(z.shape, np.uint8), h, 1.0, 0.05, 0.03)
    o.step(200)
    assert not o.read(O.U).any() and not o.read(O.V).any()
    assert np.array_equal(o.read(O.H), h)


def test_dam_break_rows_identical_and_transpose_symmetric():
    """Config 1 is one-dimensional: every DEM row evolves identically; the
    transposed grid gives the same depths with U and V exchanged."""
    z, h = dam_break_1d(20, 60)
    a = O.OracleSim(z, np.ones(z.shape, np.uint8), h, 1.0, 0.01, 0.01)
    zt, ht = dam_break_1d(20, 60, transpose=True)
    b = O.OracleSim(zt, np.ones(zt.shape, np.uint8), ht, 1.0, 0.01, 0.01)
    a.step(150)
    b.step(150)
    ha = a.read(O.H)
    assert (ha == ha[0:1, :]).all()
    assert (b.read(O.H) == ha.T).all()
    assert (b.read(O.U) == a.read(O.V).T).all()
    assert (b.read(O.V) == a.read(O.U).T).all()


def test_friction_decay_single_step():
    """Uniform flow on a flat, uniformly deep channel: only friction acts at
    an interior face, U1 = U0 (1 - D ram/2) / (1 + D ram/2) (L:194-203)."""
    R, C = 8, 8
    z = np.zeros((R, C))
    h = np.full((R, C), 2.0)
    o = O.OracleSim(z, np.ones((R, C), np.uint8), h, 1.0, 0.01, 0.05)
    U0 = np.zeros((R + 1, C))
    U0[1:R] = 0.4
    # load a uniform discharge: state write via stepping zero iterations is
    # not possible in the oracle, so test the formula on the first update
    # of a run whose U field is set through the numpy transcription instead
    r = NP.RefLayoutSim(z, h, -9999.0, 1.0, 0.01, 0.05)
    r.um[:R * C] = 0.0
    r.um[C:R * C] = 0.4
    r.umo[:] = r.um
    r.uu1_vv1()
    r.uua_vva()
    r.flux()
    D = 0.02
    uu1 = 0.4 / 2.0
    ram = G * 0.05 * 0.05 * math.sqrt(uu1 * uu1) / math.pow(2.0, 1.3333330)
    expect = ((1.00 - D * ram * 0.50) * 0.4 + (-0.0 - 0.0 - 0.0)) / (1.00 + D * ram * 0.50)
    i = 4
    got = r.u_faces()[i, 3]
    # interior face away from the walls: no convection (uniform), no gravity
    assert got == expect
    del o


def test_rain_on_still_water_exact():
    """Still water + uniform rain: h grows by D*rr*p each iteration, with p
    switching at EV_WL_min (L:627-634); the accounting matches."""
    z = np.zeros((6, 7))
    h = np.full((6, 7), 0.5)
    h[0, :] = 0.01  # below EV_WL_min: p = 1 - (INT + INF)
    h[:] = 0.5
    table = np.full(10, 6.0)   # 6 mm per 60 s
    o = O.OracleSim(z, np.ones(z.shape, np.uint8), h, 1.0, 0.5, 0.03, INT=0.45, INF=0.2,
                    LWL=0.3, EV_WL_min=0.05, forcing_mode=1)
    o.set_rain(table, 60.0)
    o.step(1)
    rr = 6.0 / (60.0 * 1000.0)
    D = 1.0
    p = 1.00 - 0.3
    assert np.all(o.read(O.H) == 0.5 - D * (0.0 / 1.0 + 0.0 / 1.0 - 0.0 - rr * p + 0.0))
    st = o.stats()
    assert abs(st["rain_in"] - D * rr * p * 42) < 1e-15
    assert abs(st["volume"] - (0.5 + D * rr * p) * 42) < 1e-12


def test_closed_walls_conserve_volume():
    z, h = random_case(11, 30, 40, holes=0.03)
    act = (z != -9999.0).astype(np.uint8)
    o = O.OracleSim(z, act, h, 1.0, 0.02, 0.02)
    v0 = o.stats()["volume"]
    o.step(400)
    st = o.stats()
    assert abs(st["volume"] + 0 - (v0 + st["clamp_gain"])) <= 1e-11 * v0


def test_treat_error_zeroes_only_exporting_faces():
    """treat_error (L:704-719): a cell that continuity 1 leaves dry (h1 < th)
    may not export; its outflow face is zeroed and the final update keeps its
    water.  One row: a thin wet cell spilling east by rule 2 with a step D
    large enough to drain it below th."""
    z = np.zeros((1, 3))
    h = np.array([[2.0e-4, 0.0, 0.0]])
    D = 40.0
    q = 0.350 * 2.0e-4 * math.sqrt(2.00 * G * 2.0e-4)      # rule 2 flux
    assert 2.0e-4 - D * q < TH                              # drained below th
    o = O.OracleSim(z, np.ones((1, 3), np.uint8), h, 1.0, D / 2, 0.0)
    o.step(1)
    assert np.array_equal(o.read(O.H), h)
    assert not o.read(O.V).any()
    # with half the step the cell stays wet and the face keeps the flux
    o2 = O.OracleSim(z, np.ones((1, 3), np.uint8), h, 1.0, D / 8, 0.0)
    o2.step(1)
    assert o2.read(O.V)[0, 1] == q
    r = NP.RefLayoutSim(z, h, -9999.0, 1.0, D / 2, 0.0)
    r.iterate()
    assert np.array_equal(r.cells(r.h), h)


def test_stream_rate_closed_forms():
    """stream_flow / evaporation_load interpolation (L:533-538): ramp a0*t for
    t <= 1, a[it] + (a[it+1]-a[it]) / (dT*(t - dT*it)) beyond, a division by
    zero on an exact breakpoint, and the piecewise-constant mode."""
    a = np.array([2.0, 4.0, 8.0])
    assert O.stream_rate(0.5, a, 60.0, 0) == 1.0
    assert O.stream_rate(1.0, a, 60.0, 0) == 2.0
    t = 75.0
    assert O.stream_rate(t, a, 60.0, 0) == 4.0 + (8.0 - 4.0) / (60.0 * (t - 60.0))
    assert math.isinf(O.stream_rate(60.0, a, 60.0, 0))
    b = np.array([3.0, 3.0])
    assert math.isnan(O.stream_rate(60.0, b, 60.0, 0))
    assert O.stream_rate(75.0, a, 60.0, 1) == 4.0
    assert O.stream_rate(1e9, a, 60.0, 1) == 8.0


def test_evaporation_entry_closed_form():
    """gpu_evaporation_calc (L:561-582) at T = 20 C, Rg = 500 W/m2."""
    T, Rg, alb, dtr = 20.0, 500.0, 0.25, 3600.0
    pw = 999.85308 + 6.32693e-2 * T - 8.523829e-3 * T ** 2 + 6.943248e-5 * T ** 3 - 3.821216e-7 * T ** 4
    lv = ((T + 273.15) / (T + 273.15 - 33.91)) ** 2 * 1.91846e6
    e = (1 - alb) * Rg / (lv * pw) * dtr * 1000.0
    assert abs(O.evap_entry(alb, T, Rg, dtr) - e) < 1e-12 * e
    assert abs(pw - 998.2) < 0.05          # ITS-90 density of water at 20 C
    assert abs(lv - 2.453e6) < 2e3         # latent heat at 20 C
    assert O.evap_entry(alb, T, -10.0, dtr) == 0.0


def test_open_edge_outflow_closed_form():
    """An open south edge behaves as dry-bed rule 1a against a bed at -inf:
    U(R,j) = 0.544 h sqrt(g h); the outflow volume is accounted."""
    z = np.zeros((4, 3))
    h = np.full((4, 3), 0.5)
    o = O.OracleSim(z, np.ones(z.shape, np.uint8), h, 2.0, 0.01, 0.02, edge_open=O.EDGE_SOUTH)
    o.step(1)
    U = o.read(O.U)
    q = 0.5440 * 0.5 * math.sqrt(G * 0.5)
    assert np.all(U[4, :] == q)
    st = o.stats()
    assert abs(st["outflow"] - 0.02 * q / 2.0 * 3 * 4.0) < 1e-14
    v = o.stats()["volume"]
    assert abs(v + st["outflow"] - 0.5 * 12 * 4.0) < 1e-12


def test_cfl_step_on_still_water():
    """Adaptive step on still water: D = min(2 dt, Cr * dx / sqrt(g h))."""
    z = np.zeros((5, 6))
    h = np.full((5, 6), 0.8)
    o = O.OracleSim(z, np.ones(z.shape, np.uint8), h, 3.0, 10.0, 0.03, cfl=0.5)
    o.step(1)
    assert o.stats()["last_D"] == 0.5 / ((0.0 + math.sqrt(G * 0.8)) / 3.0)
    o2 = O.OracleSim(z, np.ones(z.shape, np.uint8), h, 3.0, 0.01, 0.03, cfl=0.5)
    o2.step(1)
    assert o2.stats()["last_D"] == 0.02
