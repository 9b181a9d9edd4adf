"""Synthetic workloads of the benchmark configurations, built with the
library's seeded generators (include/sw2d_gen.h).  Seeds: configuration c
uses 1000 + c, ensemble grid g uses 2000 + g, serving request r 3000 + r.

Each builder returns a dict: z, h0, dx, n_cell (or None), params (keyword
arguments of sw2d.Sim), rain (table_mm, dtrain[, block_map, block]) and a
short description.
"""
import numpy as np

import sw2d


def config1(transpose=False):
    """Plain FP64 dam break: 200 x 100 cells (x along columns), dx = 1 m,
    flat bed, closed walls, n = 0.01, h = 1 for x < 100 m else 0.1,
    dt = 0.01 s (step 2*dt), 1000 iterations, no rain."""
    rows, cols = 100, 200
    z = np.zeros((rows, cols))
    h = np.where(np.arange(cols)[None, :] < 100, 1.0, 0.1) * np.ones((rows, cols))
    if transpose:
        z, h = z.T.copy(), h.T.copy()
    return dict(z=z, h0=h, dx=1.0, n_cell=None,
                params=dict(dt=0.01, manning=0.01, precision=sw2d.FP64),
                rain=None, iterations=1000,
                desc="dam break 200x100 dx=1 FP64 fixed dt=0.01 (1000 iterations)")


def catchment(size, seed, dx, relief, tilt=0.0, n_range=(0.03, 0.10)):
    z = sw2d.gen_fractal_dem(size, size, seed, 0.6, 0.0, relief)
    if tilt:
        z = sw2d.gen_tilt(z, dx, tilt, sw2d.EDGE_SOUTH)
    n = sw2d.gen_uniform((size, size), seed + 7, *n_range)
    return z, n


def config2(size=1024, precision=sw2d.FP32):
    """Single catchment: 1024^2 at dx = 5 m, fractal relief 0-50 m + 0.5%
    tilt falling to the south edge (free outflow), n ~ U(0.03, 0.10) per
    cell, dry start, 50 mm/h uniform rain for 3600 s, CFL 0.5 (step <= 2 s)."""
    z, n = catchment(size, 1002, 5.0, 50.0, tilt=0.005)
    dtrain = 300.0
    table = np.full(13, 50.0 * dtrain / 3600.0)
    return dict(z=z, h0=np.zeros_like(z), dx=5.0, n_cell=n,
                params=dict(dt=1.0, manning=0.05, cfl=0.5, edge_open=sw2d.EDGE_SOUTH,
                            forcing_mode=sw2d.FORCING_STEP, precision=precision,
                            t_end=3600.0),
                rain=(table, dtrain), iterations=None,
                desc="catchment %d^2 dx=5 fractal 0-50 m + 0.5%% tilt, open south edge, "
                     "50 mm/h, CFL 0.5, 3600 s" % size)


def config3(size=8192, precision=sw2d.FP32, reservoir=512):
    """Large grid: 8192^2 at dx = 2 m, fractal relief 0-100 m, n ~
    U(0.03, 0.10), a 10 m reservoir of 512^2 cells at the centre released at
    t = 0, rain 30 mm/h x lognormal(0, 0.5) per 64x64 block, CFL 0.5 (step
    <= 1 s), closed walls."""
    z, n = catchment(size, 1003, 2.0, 100.0)
    h = np.zeros_like(z)
    r0 = size // 2 - reservoir // 2
    h[r0:r0 + reservoir, r0:r0 + reservoir] = 10.0
    nb = (size + 63) // 64
    bmap = sw2d.gen_lognormal((nb, nb), 1003 + 11, 0.0, 0.5)
    dtrain = 3600.0
    table = np.full(48, 30.0)
    return dict(z=z, h0=h, dx=2.0, n_cell=n,
                params=dict(dt=0.5, manning=0.05, cfl=0.5, forcing_mode=sw2d.FORCING_STEP,
                            precision=precision),
                rain=(table, dtrain, bmap, 64), iterations=20000,
                desc="grid %d^2 dx=2 fractal 0-100 m, 10 m reservoir %d^2, 30 mm/h x "
                     "lognormal(0,0.5) per 64^2 block, CFL 0.5" % (size, reservoir))


def config4(n_grids=64, per_grid=4, precision=sw2d.FP32):
    """Batched ensemble: n_grids grids, side = round(exp(U(ln 256, ln 2048)))
    at dx = 5 m (grid g: seed 2000 + g), fractal relief 0-50 m + 0.5% tilt
    falling to an open south edge (recorded choice: the configuration fixes
    only the generator and dx), n ~ U(0.03, 0.10) per cell, dry start.  Each
    grid carries per_grid hyetographs of 24 five-minute intensities ~
    Gamma(k = 0.8, theta = 20 mm/h) over 2 h; CFL 0.5 (step <= 2 s).  The
    scenarios of one grid share its bed and raster on the device."""
    grids = []
    for g in range(n_grids):
        seed = 2000 + g
        u = float(sw2d.gen_uniform((1,), seed + 101, np.log(256.0), np.log(2048.0))[0])
        side = int(round(np.exp(u)))
        z, n = catchment(side, seed, 5.0, 50.0, tilt=0.005)
        hyet = [sw2d.gen_gamma((24,), seed * 16 + k, 0.8, 20.0) * (300.0 / 3600.0)
                for k in range(per_grid)]  # mm per 300 s
        grids.append(dict(z=z, n_cell=n, side=side, hyetographs=hyet))
    params = dict(dt=1.0, manning=0.05, cfl=0.5, edge_open=sw2d.EDGE_SOUTH,
                  forcing_mode=sw2d.FORCING_STEP, precision=precision, t_end=7200.0)
    return dict(grids=grids, dx=5.0, dtrain=300.0, params=params,
                desc="%d grids x %d hyetographs, sides log-uniform [256, 2048], dx=5, fractal "
                     "0-50 m + 0.5%% tilt to an open south edge, gamma(0.8, 20 mm/h) 5-min "
                     "intensities, 2 h, CFL 0.5" % (n_grids, per_grid))


def make_batch(w, precision=sw2d.FP32):
    """All scenarios of a config4 workload in one sw2d.Batch; returns
    (batch, scenario ids, cells per scenario)."""
    b = sw2d.Batch(precision)
    ids, cells = [], []
    p = dict(w["params"])
    p["precision"] = precision
    for g in w["grids"]:
        first = -1
        h0 = np.zeros_like(g["z"])
        for hy in g["hyetographs"]:
            sid = b.add(g["z"], h0, w["dx"], share_with=first, n_cell=g["n_cell"], **p)
            b.scenario(sid).set_rain(hy, w["dtrain"])
            if first < 0:
                first = sid
            ids.append(sid)
            cells.append(g["z"].size)
    b.touch()
    return b, ids, np.array(cells, dtype=np.int64)


def config5(n_requests=48, rate=50.0, precision=sw2d.FP32):
    """Serving stream: Poisson arrivals at `rate` per second (arrival times
    from seed 3000), request r (seed 3000 + r): side log-uniform in
    [128, 4096] at dx = 5 m, fractal relief 0-50 m + 0.5% tilt to an open
    south edge (recorded choice), n ~ U(0.02, 0.12) per cell, dry start,
    duration ~ U(600, 21600) s, gamma(0.8, 20 mm/h) five-minute intensities
    over the duration; CFL 0.5 (step <= 2 s)."""
    gaps = sw2d.gen_uniform((n_requests,), 3000, 0.0, 1.0)
    arrivals = np.cumsum(-np.log1p(-gaps) / rate)
    reqs = []
    for r in range(n_requests):
        seed = 3000 + r
        u = sw2d.gen_uniform((2,), seed + 101, 0.0, 1.0)
        side = int(round(np.exp(np.log(128.0) + u[0] * (np.log(4096.0) - np.log(128.0)))))
        duration = 600.0 + u[1] * (21600.0 - 600.0)
        z, n = catchment(side, seed, 5.0, 50.0, tilt=0.005, n_range=(0.02, 0.12))
        n_rain = int(np.ceil(duration / 300.0)) + 1
        rain = sw2d.gen_gamma((n_rain,), seed * 16, 0.8, 20.0) * (300.0 / 3600.0)
        reqs.append(dict(z=z, n_cell=n, side=side, duration=duration, rain=rain,
                         arrival=float(arrivals[r])))
    params = dict(dt=1.0, manning=0.05, cfl=0.5, edge_open=sw2d.EDGE_SOUTH,
                  forcing_mode=sw2d.FORCING_STEP, precision=precision)
    return dict(requests=reqs, dx=5.0, dtrain=300.0, params=params, rate=rate,
                desc="%d Poisson arrivals at %g/s, sides log-uniform [128, 4096], dx=5, "
                     "durations U(600, 21600) s, fractal 0-50 m + 0.5%% tilt to an open south "
                     "edge, n ~ U(0.02, 0.12), gamma(0.8, 20 mm/h) 5-min rain, CFL 0.5"
                     % (n_requests, rate))


def make_sim(w, **over):
    p = dict(w["params"])
    p.update(over)
    s = sw2d.Sim(w["z"], w["h0"], w["dx"], n_cell=w["n_cell"], **p)
    if w["rain"] is not None:
        s.set_rain(*w["rain"])
    return s
