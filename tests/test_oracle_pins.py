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


def run_pair(z, h0, dx, dt, n, iters, rain=None, sources=None, evap=None, **kw):
    """The C oracle and the literal transcription on the same inputs.
    rain: (table_mm, dtrain); sources: (q[n_rows][nst], dtoq, flat cells);
    evap: evaporation table in mm per dtrain (needs rain's dtrain, and turns
    evaporation on)."""
    act = (z != -9999.0).astype(np.uint8)
    if evap is not None:
        kw["evaporation_on"] = 1
    o = O.OracleSim(z, act, h0, dx, dt, n, **kw)
    r = NP.RefLayoutSim(z, h0, -9999.0, dx, dt, n, **{k: v for k, v in kw.items()
                                                     if k in ("INT", "INF", "LWL", "EV_WL_min")})
    rain_fn, np_src, np_ev = None, None, None
    if rain is not None:
        table, dtrain = rain
        o.set_rain(table, dtrain)
        rain_fn = lambda t: NP._interp(table, t, dtrain) / (dtrain * 1000.0)  # noqa: E731
    if sources is not None:
        q, dtoq, cells = sources
        o.set_sources(np.asarray(q, dtype=np.float64), dtoq, cells)
        np_src = (q, dtoq, cells)
    if evap is not None:
        o.set_evaporation(evap)
        np_ev = (evap, rain[1])
    for _ in range(iters):
        o.step(1)
        r.iterate(rain_fn, sources=np_src, evap=np_ev)
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
    """Still water + uniform rain: h grows by D*rr*p each iteration (L:654)
    and the accounting matches.  The switch of p at EV_WL_min is pinned by
    test_evaporation_and_runoff_fraction_on_still_water."""
    z = np.zeros((6, 7))
    h = np.full((6, 7), 0.5)
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


def test_oracle_matches_literal_with_sources_and_evaporation():
    """Point sources (two on one cell: the later one wins, L:540-545) and
    evaporation (L:589-601, switched by EV_WL_min at L:627-634) with rain:
    the C oracle and the literal transcription agree bitwise."""
    z, h = random_case(17, 12, 15, holes=0.04)
    R, Cc = z.shape
    cells = [3 * Cc + 4, 7 * Cc + 11, 3 * Cc + 4, 10 * Cc + 2]
    q = [[0.30, -0.10, 0.05, 0.20], [0.10, -0.25, 0.40, 0.15], [0.00, -0.05, 0.60, 0.25],
         [0.20, -0.15, 0.10, 0.05], [0.35, -0.20, 0.30, 0.10], [0.15, -0.05, 0.20, 0.30]]
    rain = (np.array([3.0, 5.0, 4.0, 1.0]), 7.0)
    evap = np.array([1.5, 2.5, 0.5, 2.0])
    o, r = run_pair(z, h, 2.0, 0.03, 0.05, 300, rain=rain, sources=(q, 5.0, cells), evap=evap,
                    INT=0.25, INF=0.15, LWL=0.1, EV_WL_min=0.2)
    assert_same(o, r)
    # the sources and the evaporation did act
    o2, r2 = run_pair(z, h, 2.0, 0.03, 0.05, 300, rain=rain, INT=0.25, INF=0.15, LWL=0.1,
                      EV_WL_min=0.2)
    assert not np.array_equal(o.read(O.H), o2.read(O.H))
    st = o.stats()
    assert st["source_in"] != 0.0 and st["evap_out"] > 0.0


def test_point_source_on_still_water_closed_form():
    """Still water (flat surface, no flow) with point sources: one iteration
    raises each source cell by exactly D q / (dx dy) (L:540, L:654); two
    sources on one cell do not add up, the later one wins (L:545 assigns);
    a cell index outside the grid is ignored; the other cells keep their
    depth bitwise.  Dyadic values make every expected depth exact."""
    R, Cc, dx = 5, 6, 2.0
    z = np.zeros((R, Cc))
    h = np.full((R, Cc), 0.5)
    c1, c2 = 1 * Cc + 2, 3 * Cc + 5
    q = np.array([[0.25, 0.5, 0.125, 3.0]])       # m3/s; the last source is outside
    cells = [c1, c2, c1, R * Cc]
    o = O.OracleSim(z, np.ones(z.shape, np.uint8), h, dx, 0.5, 0.03, forcing_mode=1)
    o.set_sources(q, 10.0, cells)
    o.step(1)
    D = 1.0
    got = o.read(O.H).ravel()
    expect = np.full(R * Cc, 0.5)
    expect[c1] = 0.5 + D * 0.125 / (dx * dx)     # 0.53125: the later source wins
    expect[c2] = 0.5 + D * 0.5 / (dx * dx)       # 0.625
    assert np.array_equal(got, expect)
    assert o.stats()["source_in"] == D * (0.125 + 0.5)
    # reference forcing formula: the ramp q[0] * t for t <= 1 s (L:533-534),
    # so the first continuity (t = 0) adds nothing and the second (t = dt)
    # adds D q t / (dx dy)
    o2 = O.OracleSim(z, np.ones(z.shape, np.uint8), h, dx, 0.25, 0.03, forcing_mode=0)
    o2.set_sources(np.array([[0.25], [0.25]]), 10.0, [c2])
    o2.step(1)
    got2 = o2.read(O.H).ravel()
    assert got2[c2] == 0.5 + 0.5 * (0.25 * 0.25) / (dx * dx)    # 0.5078125
    assert np.count_nonzero(got2 != 0.5) == 1


def test_evaporation_and_runoff_fraction_on_still_water():
    """Still water with rain and evaporation (L:627-660): a cell whose depth
    at the start of the iteration is above EV_WL_min loses D ev and gains
    D rr (1 - LWL); a cell at or below it gains D rr (1 - (INT + INF)) and
    does not evaporate.  The free surface is flat (bed = 0.5 - depth) so no
    face carries flow."""
    depths = np.array([0.5, 0.03125, 0.0625, 0.25, 0.0625 + 2.0 ** -20])
    EVmin = 0.0625
    h = np.repeat(depths[:, None], 4, axis=1)
    z = 0.5 - h
    INT, INF, LWL = 0.25, 0.125, 0.375
    dtrain = 60.0
    rain_mm, ev_mm = 6.0, 1.5
    o = O.OracleSim(z, np.ones(z.shape, np.uint8), h, 1.0, 0.5, 0.03, INT=INT, INF=INF, LWL=LWL,
                    EV_WL_min=EVmin, evaporation_on=1, forcing_mode=1)
    o.set_rain(np.full(4, rain_mm), dtrain)
    o.set_evaporation(np.full(4, ev_mm))
    o.step(1)
    D = 1.0
    rr = rain_mm / (dtrain * 1000.0)
    ev = ev_mm / (dtrain * 1000.0)
    hi = depths > EVmin
    assert hi.tolist() == [True, False, False, True, True]
    expect = np.where(hi, depths - D * ev + D * rr * (1.0 - LWL), depths + D * rr * (1.0 - (INT + INF)))
    got = o.read(O.H)
    assert np.allclose(got, expect[:, None], rtol=0, atol=2e-16), np.abs(got - expect[:, None]).max()
    # a wrong switch (h >= EV_WL_min) or wrong fraction would move a row by
    # at least D rr * 0.25 = 2.5e-5
    assert not o.read(O.U).any() and not o.read(O.V).any()
    st = o.stats()
    assert abs(st["evap_out"] - D * ev * 4 * hi.sum()) < 1e-18
    p = np.where(hi, 1.0 - LWL, 1.0 - (INT + INF))
    assert abs(st["rain_in"] - D * rr * 4 * p.sum()) < 1e-18
    # evaporation off (L:1283): the table is ignored
    o2 = O.OracleSim(z, np.ones(z.shape, np.uint8), h, 1.0, 0.5, 0.03, INT=INT, INF=INF, LWL=LWL,
                     EV_WL_min=EVmin, evaporation_on=0, forcing_mode=1)
    o2.set_rain(np.full(4, rain_mm), dtrain)
    o2.set_evaporation(np.full(4, ev_mm))
    o2.step(1)
    expect2 = depths + D * rr * p
    assert np.allclose(o2.read(O.H), expect2[:, None], rtol=0, atol=2e-16)
    assert o2.stats()["evap_out"] == 0.0


@pytest.mark.parametrize("nval", [0.03, 0.0731, 0.11])
def test_uniform_raster_equals_scalar_manning(nval):
    """A uniform Manning raster gives the scalar manning_coef run bitwise:
    the face value sqrt(0.5 (n^2 + n^2)) is n exactly in binary floating
    point (sqrt of a correctly rounded square returns |n|)."""
    z, h = random_case(23, 14, 19, holes=0.05)
    act = (z != -9999.0).astype(np.uint8)
    a = O.OracleSim(z, act, h, 1.5, 0.02, nval, edge_open=O.EDGE_SOUTH)
    b = O.OracleSim(z, act, h, 1.5, 0.02, 0.5, edge_open=O.EDGE_SOUTH,
                    n_cell=np.full(z.shape, nval))
    a.step(200)
    b.step(200)
    for f in (O.H, O.U, O.V):
        assert np.array_equal(a.read(f), b.read(f))


def _friction_face(n_a, n_b, along_rows):
    """One face with only friction acting: a column (along_rows) or a row of
    8 cells, flat bed, depth 2, discharge 0.4 on every interior face; the
    middle face lies between cells 3 and 4.  Returns (U or V at that face
    after one iteration, the step D, uu1 at the face)."""
    n = np.full(8, 0.05)
    n[3], n[4] = n_a, n_b
    if along_rows:
        shape, ncell = (8, 1), n[:, None]
    else:
        shape, ncell = (1, 8), n[None, :]
    z = np.zeros(shape)
    h = np.full(shape, 2.0)
    o = O.OracleSim(z, np.ones(shape, np.uint8), h, 1.0, 0.01, 0.5, n_cell=ncell)
    U = np.zeros((shape[0] + 1, shape[1]))
    V = np.zeros((shape[0], shape[1] + 1))
    if along_rows:
        U[1:8, 0] = 0.4
    else:
        V[0, 1:8] = 0.4
    o.set_state(h, U, V)
    o.step(1)
    got = o.read(O.U)[4, 0] if along_rows else o.read(O.V)[0, 4]
    return got, 0.02, 0.4 / 2.0


@pytest.mark.parametrize("along_rows", [True, False])
def test_manning_raster_face_rule(along_rows):
    """Face roughness of the raster: n_face^2 = (n_a^2 + n_b^2) / 2 in the
    implicit friction update U1 = U0 (1 - D ram/2) / (1 + D ram/2), ram =
    g n_face^2 |u| / hf^1.3333330 (L:194-203 for U, L:311-320 for V).  The
    expected value is formed from the mean of the squares directly; the
    arithmetic mean of n, or either cell's n alone, miss it by far more than
    the tolerance."""
    na, nb = 0.03, 0.11
    got, D, u = _friction_face(na, nb, along_rows)

    def decay(nf2):
        ram = G * nf2 * abs(u) / math.pow(2.0, 1.3333330)
        return (1.0 - D * ram / 2) * 0.4 / (1.0 + D * ram / 2)

    expect = decay(0.5 * (na * na + nb * nb))
    assert abs(got - expect) <= 1e-15, (got, expect)
    for wrong in (((na + nb) / 2) ** 2, na * na, nb * nb):
        assert abs(decay(wrong) - expect) > 1e-7
    # swapping the two cells gives the same face value
    got2, _, _ = _friction_face(nb, na, along_rows)
    assert abs(got2 - expect) <= 1e-15


def test_rain_block_map_on_still_water():
    """Spatial rain (block multiplier map x table): on still water every cell
    grows by D rr m p with m its block's multiplier; the map is row-major
    over ceil(rows/b) x ceil(cols/b) blocks of b x b cells, the last blocks
    ragged.  The expected field is built by repeating the map (no index
    formula), and a map of ones equals no map bitwise."""
    R, Cc, b = 11, 13, 4
    br, bc = (R + b - 1) // b, (Cc + b - 1) // b
    bmap = (np.arange(br * bc, dtype=np.float64) + 1.0) * 0.25   # 0.25, 0.5, ... distinct
    z = np.zeros((R, Cc))
    h = np.full((R, Cc), 0.5)
    LWL = 0.25
    table = np.full(4, 7.5)
    dtrain = 60.0
    o = O.OracleSim(z, np.ones(z.shape, np.uint8), h, 1.0, 0.5, 0.03, LWL=LWL, forcing_mode=1)
    o.set_rain(table, dtrain, bmap, b)
    o.step(1)
    D = 1.0
    rr = 7.5 / (dtrain * 1000.0)
    m = np.repeat(np.repeat(bmap.reshape(br, bc), b, axis=0), b, axis=1)[:R, :Cc]
    expect = 0.5 + D * rr * m * (1.0 - LWL)
    got = o.read(O.H)
    assert np.allclose(got, expect, rtol=0, atol=2e-16), np.abs(got - expect).max()
    # the transposed map would put other multipliers on at least one cell
    mt = np.repeat(np.repeat(bmap.reshape(bc, br).T, b, axis=0), b, axis=1)[:R, :Cc]
    assert np.abs(mt - m).max() > 0
    assert abs(o.stats()["rain_in"] - D * rr * (1.0 - LWL) * m.sum()) < 1e-15
    # a map of ones is the uniform rain of the table
    zr, hr = random_case(29, R, Cc, holes=0.05)
    act = (zr != -9999.0).astype(np.uint8)
    u = O.OracleSim(zr, act, hr, 1.0, 0.05, 0.03, LWL=LWL)
    w = O.OracleSim(zr, act, hr, 1.0, 0.05, 0.03, LWL=LWL)
    u.set_rain(np.array([4.0, 9.0, 2.0]), 5.0)
    w.set_rain(np.array([4.0, 9.0, 2.0]), 5.0, np.ones(br * bc), b)
    u.step(150)
    w.step(150)
    for f in (O.H, O.U, O.V):
        assert np.array_equal(u.read(f), w.read(f))


def _pow_args(n=6000, seed=5):
    import random
    rnd = random.Random(seed)
    xs = [10 ** rnd.uniform(-4, 3) for _ in range(n)]
    # the reduction boundaries, powers of two, the threshold, odd cases
    xs += [1.0, 2.0, 0.5, 0.25, 1e-4, 1.0000000000000002, 0.9999999999999999,
           0.70710678118654746, 0.70710678118654757, 1.4142135623730949, 1.4142135623730951, 3.0, 10.0,
           100.0, 999.0]
    return xs


def test_friction_pow_error_against_60_digit_evaluation():
    """The oracle's friction power (orc_pow_fric, shared with the device FP64
    step) against x^1.3333330 evaluated with 60-digit decimal arithmetic:
    within 0.65 ulp everywhere on [1e-4, 1e3] (the CUDA pow the reference
    calls guarantees 2 ulp); a dropped series term or a wrong ln 2 split
    gives errors of many ulps."""
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    Y = Decimal(1.3333330)
    worst = 0.0
    for x in _pow_args():
        r = (Y * Decimal(x).ln()).exp()
        err = abs(Decimal(O.pow_fric(x)) - r) / Decimal(math.ulp(float(r)))
        worst = max(worst, float(err))
    assert worst <= 0.65, worst
    assert O.pow_fric(1.0) == 1.0 and O.pow_fric(2.0) == math.pow(2.0, 1.3333330)


def test_friction_pow_agrees_with_glibc_within_one_ulp():
    n_diff = 0
    xs = _pow_args(seed=6)
    for x in xs:
        a, g = O.pow_fric(x), math.pow(x, 1.3333330)
        assert abs(a - g) <= math.ulp(g), (x, a, g)
        n_diff += a != g
    assert n_diff <= 0.05 * len(xs)
