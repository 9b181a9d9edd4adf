"""ctypes binding of libsw2d.so (the C ABI in include/sw2d.h).

Python is only a harness here: tests, the smoke check and bench.py drive
the C ABI through this module.  Loading fails loudly when the HIP library
has not been built; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsw2d.so")

OK, E_INVALID, E_DEVICE, E_IO, E_MEMORY, E_STATE, E_NONFINITE = 0, -1, -2, -3, -4, -5, -6
EDGE_NORTH, EDGE_SOUTH, EDGE_WEST, EDGE_EAST = 1, 2, 4, 8
FP64, FP32 = 0, 1
FORCING_REFERENCE, FORCING_STEP = 0, 1
F_H, F_U, F_V, F_UU, F_VV, F_Z = range(6)
CASE_DEM, CASE_IC, CASE_RAIN, CASE_Q, CASE_SRC_XY, CASE_OUT_XY, CASE_TEMP, CASE_RG, CASE_EVAPO = range(9)

EXPORTS = [
    "sw2d_params_default", "sw2d_create", "sw2d_destroy", "sw2d_set_rain",
    "sw2d_set_sources", "sw2d_set_evaporation", "sw2d_set_dt_schedule",
    "sw2d_record_dt", "sw2d_get_dt_history", "sw2d_step", "sw2d_sync",
    "sw2d_read", "sw2d_read_noalloc", "sw2d_output_begin", "sw2d_output_wait", "sw2d_write_state", "sw2d_get_stats", "sw2d_snapshot",
    "sw2d_restore", "sw2d_footprint_bytes", "sw2d_last_error",
    "sw2d_evaporation_table", "sw2d_forcing_value", "sw2d_friction_pow", "sw2d_case_load",
    "sw2d_case_free", "sw2d_case_get_info", "sw2d_case_get_array",
    "sw2d_case_source_cells", "sw2d_case_create", "sw2d_write_vtk",
    "sw2d_run_db", "sw2d_profile_begin", "sw2d_profile_end",
    "sw2d_gen_fractal_dem", "sw2d_gen_tilt", "sw2d_gen_uniform",
    "sw2d_gen_lognormal", "sw2d_gen_gamma", "sw2d_batch_create", "sw2d_batch_destroy",
    "sw2d_batch_add", "sw2d_batch_scenario", "sw2d_batch_touch", "sw2d_batch_remove", "sw2d_batch_detach", "sw2d_batch_attach",
    "sw2d_batch_set_tile_rows", "sw2d_batch_autotune", "sw2d_batch_tile_rows",
    "sw2d_batch_step", "sw2d_batch_sync", "sw2d_batch_status", "sw2d_batch_idle", "sw2d_device_memory",
    "sw2d_server_create", "sw2d_server_destroy", "sw2d_server_submit", "sw2d_server_set_active_cells",
    "sw2d_server_advance",
    "sw2d_server_info", "sw2d_server_read", "sw2d_server_stats", "sw2d_server_release",
]

REQ_QUEUED, REQ_RUNNING, REQ_DONE = 0, 1, 2


class Grid(C.Structure):
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("dx", C.c_double),
                ("dy", C.c_double), ("nodata", C.c_double),
                ("z", C.POINTER(C.c_double)), ("h0", C.POINTER(C.c_double)),
                ("n_cell", C.POINTER(C.c_double))]


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("g", C.c_double), ("th", C.c_double),
                ("manning", C.c_double), ("INT", C.c_double), ("INF", C.c_double),
                ("LWL", C.c_double), ("EV_WL_min", C.c_double),
                ("evaporation_on", C.c_int), ("cfl", C.c_double),
                ("edge_open", C.c_int), ("forcing_mode", C.c_int),
                ("precision", C.c_int), ("t0", C.c_double), ("t_end", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("t", C.c_double), ("iterations", C.c_int64), ("last_D", C.c_double),
                ("done", C.c_int), ("volume", C.c_double), ("rain_in", C.c_double),
                ("source_in", C.c_double), ("evap_out", C.c_double),
                ("outflow", C.c_double), ("clamp_gain", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int), ("rows", C.c_int), ("cols", C.c_int),
                ("submit_s", C.c_double), ("admit_s", C.c_double), ("finish_s", C.c_double),
                ("iterations", C.c_int64), ("start_s", C.c_double)]


class CaseInfo(C.Structure):
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("xll", C.c_double),
                ("yll", C.c_double), ("cellsize", C.c_double), ("nodata", C.c_double),
                ("tstart", C.c_double * 4), ("tend", C.c_double * 4),
                ("dt", C.c_double), ("dpout", C.c_double), ("dkout", C.c_double),
                ("dtoq", C.c_double), ("evaporation_on", C.c_int),
                ("EV_WL_min", C.c_double), ("INT", C.c_double), ("INF", C.c_double),
                ("LWL", C.c_double), ("manning_coef", C.c_double),
                ("out_velocity_x", C.c_int), ("out_velocity_y", C.c_int),
                ("out_elevation", C.c_int), ("out_depth", C.c_int),
                ("out_outlet_on", C.c_int), ("nst", C.c_int), ("n_q_rows", C.c_int),
                ("n_rain", C.c_int), ("n_out", C.c_int), ("dtrain", C.c_double),
                ("albedo", C.c_double), ("time0", C.c_double), ("timmax", C.c_double)]


_lib = None


def lib():
    """Load libsw2d.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsw2d.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        dp, vp = C.POINTER(C.c_double), C.c_void_p
        L.sw2d_params_default.argtypes = [C.POINTER(Params)]
        L.sw2d_create.argtypes = [C.POINTER(Grid), C.POINTER(Params), C.POINTER(vp)]
        L.sw2d_destroy.argtypes = [vp]
        L.sw2d_set_rain.argtypes = [vp, dp, C.c_int, C.c_double, dp, C.c_int]
        L.sw2d_set_sources.argtypes = [vp, dp, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int64)]
        L.sw2d_set_evaporation.argtypes = [vp, dp, C.c_int]
        L.sw2d_set_dt_schedule.argtypes = [vp, dp, C.c_int]
        L.sw2d_record_dt.argtypes = [vp, C.c_int]
        L.sw2d_get_dt_history.argtypes = [vp, dp, C.c_int, C.POINTER(C.c_int)]
        L.sw2d_step.argtypes = [vp, C.c_int64]
        L.sw2d_sync.argtypes = [vp]
        L.sw2d_read.argtypes = [vp, C.c_int, dp]
        L.sw2d_read_noalloc.argtypes = [vp, C.c_int, dp]
        L.sw2d_output_begin.argtypes = [vp, C.POINTER(C.c_int)]
        L.sw2d_output_wait.argtypes = [vp, C.c_int, C.POINTER(dp), C.POINTER(dp), C.POINTER(dp)]
        L.sw2d_write_state.argtypes = [vp, dp, dp, dp]
        L.sw2d_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.sw2d_snapshot.argtypes = [vp, vp, C.POINTER(C.c_size_t)]
        L.sw2d_restore.argtypes = [vp, vp, C.c_size_t]
        L.sw2d_footprint_bytes.restype = C.c_size_t
        L.sw2d_footprint_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.sw2d_last_error.restype = C.c_char_p
        L.sw2d_evaporation_table.argtypes = [C.c_double, dp, dp, C.c_int, C.c_double, dp]
        L.sw2d_friction_pow.restype = C.c_double
        L.sw2d_friction_pow.argtypes = [C.c_double]
        L.sw2d_forcing_value.restype = C.c_double
        L.sw2d_forcing_value.argtypes = [C.c_double, dp, C.c_int, C.c_double, C.c_int]
        L.sw2d_case_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.sw2d_case_free.argtypes = [vp]
        L.sw2d_case_get_info.argtypes = [vp, C.POINTER(CaseInfo)]
        L.sw2d_case_get_array.argtypes = [vp, C.c_int, dp, C.c_size_t, C.POINTER(C.c_size_t)]
        L.sw2d_case_source_cells.argtypes = [vp, C.POINTER(C.c_int64), C.c_int]
        L.sw2d_case_create.argtypes = [vp, C.POINTER(Params), C.POINTER(vp)]
        L.sw2d_write_vtk.argtypes = [C.c_char_p, vp, dp, dp, dp, dp]
        L.sw2d_run_db.argtypes = [C.c_char_p, C.POINTER(Params)]
        L.sw2d_profile_begin.argtypes = [vp, C.c_int]
        L.sw2d_profile_end.argtypes = [vp, dp, C.POINTER(C.c_int64)]
        L.sw2d_gen_fractal_dem.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_double, C.c_double,
                                           C.c_double, dp]
        L.sw2d_gen_tilt.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, dp]
        for f in ("sw2d_gen_uniform", "sw2d_gen_lognormal", "sw2d_gen_gamma"):
            getattr(L, f).argtypes = [C.c_size_t, C.c_uint64, C.c_double, C.c_double, dp]
        ip = C.POINTER(C.c_int)
        L.sw2d_batch_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.sw2d_batch_destroy.argtypes = [vp]
        L.sw2d_batch_add.argtypes = [vp, C.POINTER(Grid), C.POINTER(Params), C.c_int, ip]
        L.sw2d_batch_scenario.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.sw2d_batch_touch.argtypes = [vp]
        L.sw2d_batch_remove.argtypes = [vp, C.c_int]
        L.sw2d_batch_detach.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.sw2d_batch_attach.argtypes = [vp, vp, ip]
        L.sw2d_batch_set_tile_rows.argtypes = [vp, C.c_int, C.c_int]
        L.sw2d_batch_autotune.argtypes = [vp, C.c_int]
        L.sw2d_batch_tile_rows.argtypes = [vp, C.c_int, ip]
        L.sw2d_batch_step.argtypes = [vp, C.c_int64]
        L.sw2d_batch_sync.argtypes = [vp]
        L.sw2d_batch_idle.argtypes = [vp, C.POINTER(C.c_int)]
        L.sw2d_device_memory.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.sw2d_batch_status.argtypes = [vp, ip, C.POINTER(C.c_int64), C.c_int, ip]
        L.sw2d_server_create.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(vp)]
        L.sw2d_server_destroy.argtypes = [vp]
        L.sw2d_server_submit.argtypes = [vp, C.POINTER(Grid), C.POINTER(Params), dp, C.c_int,
                                         C.c_double, ip]
        L.sw2d_server_advance.argtypes = [vp, C.c_int64, ip, ip]
        L.sw2d_server_set_active_cells.argtypes = [vp, C.c_int64]
        L.sw2d_server_info.argtypes = [vp, C.c_int, C.POINTER(RequestInfo)]
        L.sw2d_server_read.argtypes = [vp, C.c_int, C.c_int, dp]
        L.sw2d_server_stats.argtypes = [vp, C.c_int, C.POINTER(Stats)]
        L.sw2d_server_release.argtypes = [vp, C.c_int]
        _lib = L
    return _lib


class SW2DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sw2d error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc != 0:
        raise SW2DError(rc, lib().sw2d_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params(**kw):
    p = Params()
    lib().sw2d_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Sim:
    """One device context (a single grid)."""

    def __init__(self, z, h0, dx, *, dy=None, nodata=-9999.0, n_cell=None, **params):
        self.rows, self.cols = z.shape
        self._z = np.ascontiguousarray(z, dtype=np.float64)
        self._h = np.ascontiguousarray(h0, dtype=np.float64)
        g = Grid(self.rows, self.cols, float(dx), float(dy if dy is not None else dx),
                 float(nodata), _dp(self._z), _dp(self._h), None)
        if n_cell is not None:
            self._n = np.ascontiguousarray(n_cell, dtype=np.float64)
            g.n_cell = _dp(self._n)
        p = default_params(**params)
        self.params = p
        h = C.c_void_p()
        check(lib().sw2d_create(C.byref(g), C.byref(p), C.byref(h)))
        self.h_ = h

    owned = True

    @classmethod
    def from_handle(cls, handle, rows, cols, owned=True):
        s = cls.__new__(cls)
        s.h_, s.rows, s.cols, s.owned = handle, rows, cols, owned
        return s

    def close(self):
        if getattr(self, "h_", None):
            if self.owned:
                lib().sw2d_destroy(self.h_)
            self.h_ = None

    def __del__(self):
        self.close()

    def set_rain(self, table_mm, dtrain, block_map=None, block=0):
        t = np.ascontiguousarray(table_mm, dtype=np.float64)
        bm = None
        if block_map is not None:
            self._bm = np.ascontiguousarray(block_map, dtype=np.float64)
            bm = _dp(self._bm)
        check(lib().sw2d_set_rain(self.h_, _dp(t), len(t), float(dtrain), bm, int(block)))

    def set_sources(self, q, dtoq, cell_index):
        q = np.ascontiguousarray(q, dtype=np.float64)
        ci = np.ascontiguousarray(cell_index, dtype=np.int64)
        check(lib().sw2d_set_sources(self.h_, _dp(q), q.shape[0], q.shape[1], float(dtoq),
                                     ci.ctypes.data_as(C.POINTER(C.c_int64))))

    def set_evaporation(self, evapo):
        e = np.ascontiguousarray(evapo, dtype=np.float64)
        check(lib().sw2d_set_evaporation(self.h_, _dp(e), len(e)))

    def set_dt_schedule(self, D):
        d = np.ascontiguousarray(D, dtype=np.float64)
        check(lib().sw2d_set_dt_schedule(self.h_, _dp(d), len(d)))

    def record_dt(self, capacity):
        check(lib().sw2d_record_dt(self.h_, int(capacity)))

    def dt_history(self, capacity):
        out = np.zeros(capacity)
        n = C.c_int()
        check(lib().sw2d_get_dt_history(self.h_, _dp(out), capacity, C.byref(n)))
        return out[:n.value]

    def step(self, n):
        check(lib().sw2d_step(self.h_, int(n)))

    def sync(self):
        check(lib().sw2d_sync(self.h_))

    def profile_begin(self, max_iters):
        check(lib().sw2d_profile_begin(self.h_, int(max_iters)))

    def profile_end(self):
        """Summed ms of [control kernel, step kernel, 0] and the profiled iteration count."""
        ms = np.zeros(3)
        n = C.c_int64()
        check(lib().sw2d_profile_end(self.h_, _dp(ms), C.byref(n)))
        return ms, n.value

    def read(self, field):
        R, Cc = self.rows, self.cols
        shape = {F_H: (R, Cc), F_UU: (R, Cc), F_VV: (R, Cc), F_Z: (R, Cc),
                 F_U: (R + 1, Cc), F_V: (R, Cc + 1)}[field]
        out = np.empty(shape)
        check(lib().sw2d_read(self.h_, field, _dp(out)))
        return out

    def output_begin(self):
        """sw2d_output_begin: snapshot (h, vv, uu) asynchronously; the slot."""
        k = C.c_int()
        check(lib().sw2d_output_begin(self.h_, C.byref(k)))
        return k.value

    def output_wait(self, slot):
        """sw2d_output_wait: copies of (h, vv, uu) of a snapshot."""
        ptrs = [C.POINTER(C.c_double)() for _ in range(3)]
        check(lib().sw2d_output_wait(self.h_, slot, *[C.byref(q) for q in ptrs]))
        n = self.rows * self.cols
        return [np.ctypeslib.as_array(q, shape=(n,)).reshape(self.rows, self.cols).copy() for q in ptrs]

    def read_noalloc(self, field):
        """sw2d_read_noalloc: h, U, V or z through the library's fixed device
        staging array (no device allocation)."""
        R, Cc = self.rows, self.cols
        shape = {F_H: (R, Cc), F_UU: (R, Cc), F_VV: (R, Cc), F_Z: (R, Cc),
                 F_U: (R + 1, Cc), F_V: (R, Cc + 1)}[field]
        out = np.empty(shape)
        check(lib().sw2d_read_noalloc(self.h_, field, _dp(out)))
        return out

    def write_state(self, h, U, V):
        h = np.ascontiguousarray(h, dtype=np.float64)
        U = np.ascontiguousarray(U, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        check(lib().sw2d_write_state(self.h_, _dp(h), _dp(U), _dp(V)))

    def stats(self):
        s = Stats()
        check(lib().sw2d_get_stats(self.h_, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def snapshot(self):
        n = C.c_size_t()
        check(lib().sw2d_snapshot(self.h_, None, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().sw2d_snapshot(self.h_, buf, C.byref(n)))
        return bytes(buf.raw[:n.value])

    def restore(self, blob):
        buf = C.create_string_buffer(blob, len(blob))
        check(lib().sw2d_restore(self.h_, buf, len(blob)))


class Case:
    """A parsed reference case directory (db/<case>)."""

    def __init__(self, case_dir):
        h = C.c_void_p()
        check(lib().sw2d_case_load(case_dir.encode(), C.byref(h)))
        self.h_ = h
        info = CaseInfo()
        check(lib().sw2d_case_get_info(h, C.byref(info)))
        self.info = info

    def __del__(self):
        if getattr(self, "h_", None):
            lib().sw2d_case_free(self.h_)
            self.h_ = None

    def array(self, which):
        n = C.c_size_t()
        check(lib().sw2d_case_get_array(self.h_, which, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        check(lib().sw2d_case_get_array(self.h_, which, _dp(out), n.value, C.byref(n)))
        return out

    def source_cells(self):
        n = self.info.nst
        out = np.zeros(max(n, 1), dtype=np.int64)
        check(lib().sw2d_case_source_cells(self.h_, out.ctypes.data_as(C.POINTER(C.c_int64)), n))
        return out[:n]

    def create(self, **overrides):
        p = default_params(**overrides) if overrides else None
        h = C.c_void_p()
        check(lib().sw2d_case_create(self.h_, C.byref(p) if p is not None else None, C.byref(h)))
        return Sim.from_handle(h, self.info.rows, self.info.cols)

    def write_vtk(self, path, depth, vv, uu, z):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (depth, vv, uu, z)]
        check(lib().sw2d_write_vtk(path.encode(), self.h_, *[_dp(a) for a in arrs]))


def run_db(db_root, **overrides):
    p = default_params(**overrides)
    check(lib().sw2d_run_db(db_root.encode(), C.byref(p)))


def footprint_bytes(rows, cols, precision=FP64, has_n=False):
    return lib().sw2d_footprint_bytes(rows, cols, precision, int(bool(has_n)))


def device_memory():
    """(free, total) bytes of the current device."""
    f, t = C.c_size_t(), C.c_size_t()
    check(lib().sw2d_device_memory(C.byref(f), C.byref(t)))
    return f.value, t.value


def friction_pow(x):
    """x^1.3333330 as the FP64 step evaluates it (sw2d_friction_pow)."""
    return lib().sw2d_friction_pow(float(x))


def forcing_value(time, table, dT, mode=0):
    t = np.ascontiguousarray(table, dtype=np.float64)
    return lib().sw2d_forcing_value(time, _dp(t), len(t), dT, mode)


def evaporation_table(albedo, T, Rg, dtrain):
    T = np.ascontiguousarray(T, dtype=np.float64)
    Rg = np.ascontiguousarray(Rg, dtype=np.float64)
    out = np.zeros(len(T))
    check(lib().sw2d_evaporation_table(albedo, _dp(T), _dp(Rg), len(T), dtrain, _dp(out)))
    return out


# ---- seeded synthetic inputs (include/sw2d_gen.h) ----------------------
def gen_fractal_dem(rows, cols, seed, roughness, zmin, zmax):
    out = np.empty((rows, cols))
    check(lib().sw2d_gen_fractal_dem(rows, cols, seed, roughness, zmin, zmax, _dp(out)))
    return out


def gen_tilt(z, dx, slope, edge):
    z = np.ascontiguousarray(z, dtype=np.float64)
    check(lib().sw2d_gen_tilt(z.shape[0], z.shape[1], dx, slope, edge, _dp(z)))
    return z


def _gen(name, shape, seed, a, b):
    out = np.empty(shape)
    check(getattr(lib(), name)(out.size, seed, a, b, _dp(out)))
    return out


def gen_uniform(shape, seed, lo, hi):
    return _gen("sw2d_gen_uniform", shape, seed, lo, hi)


def gen_lognormal(shape, seed, mu, sigma):
    return _gen("sw2d_gen_lognormal", shape, seed, mu, sigma)


def gen_gamma(shape, seed, k, theta):
    return _gen("sw2d_gen_gamma", shape, seed, k, theta)


class Batch:
    """Batched ensemble of scenarios (sw2d_batch_* in include/sw2d.h)."""

    def __init__(self, precision=FP32):
        h = C.c_void_p()
        check(lib().sw2d_batch_create(precision, C.byref(h)))
        self.h_ = h
        self.precision = precision
        self.shapes = []

    def close(self):
        if getattr(self, "h_", None):
            lib().sw2d_batch_destroy(self.h_)
            self.h_ = None

    def __del__(self):
        self.close()

    def add(self, z, h0, dx, *, share_with=-1, dy=None, nodata=-9999.0, n_cell=None, **params):
        z = np.ascontiguousarray(z, dtype=np.float64)
        h0 = np.ascontiguousarray(h0, dtype=np.float64)
        g = Grid(z.shape[0], z.shape[1], float(dx), float(dy if dy is not None else dx),
                 float(nodata), _dp(z), _dp(h0), None)
        if n_cell is not None:
            n_cell = np.ascontiguousarray(n_cell, dtype=np.float64)
            g.n_cell = _dp(n_cell)
        params.setdefault("precision", self.precision)
        p = default_params(**params)
        out = C.c_int()
        check(lib().sw2d_batch_add(self.h_, C.byref(g), C.byref(p), int(share_with), C.byref(out)))
        while len(self.shapes) <= out.value:
            self.shapes.append(None)
        self.shapes[out.value] = z.shape
        return out.value

    def scenario(self, sid):
        """Non-owning handle of one scenario's context."""
        h = C.c_void_p()
        check(lib().sw2d_batch_scenario(self.h_, sid, C.byref(h)))
        r, c = self.shapes[sid]
        return Sim.from_handle(h, r, c, owned=False)

    def touch(self):
        check(lib().sw2d_batch_touch(self.h_))

    def remove(self, sid):
        check(lib().sw2d_batch_remove(self.h_, sid))

    def detach(self, sid):
        """Takes the scenario out of the batch; returns its context as an
        owning handle (closing it destroys the context)."""
        h = C.c_void_p()
        check(lib().sw2d_batch_detach(self.h_, sid, C.byref(h)))
        r, c = self.shapes[sid]
        return Sim.from_handle(h, r, c, owned=True)

    def attach(self, sim):
        """Adds an existing context (a Sim that owns it) as a scenario; the
        batch owns the context from then on.  Returns the scenario id."""
        out = C.c_int()
        check(lib().sw2d_batch_attach(self.h_, sim.h_, C.byref(out)))
        sim.owned = False
        while len(self.shapes) <= out.value:
            self.shapes.append(None)
        self.shapes[out.value] = (sim.rows, sim.cols)
        return out.value

    def set_tile_rows(self, sid, rows):
        check(lib().sw2d_batch_set_tile_rows(self.h_, sid, rows))

    def tile_rows(self, sid):
        r = C.c_int()
        check(lib().sw2d_batch_tile_rows(self.h_, sid, C.byref(r)))
        return r.value

    def autotune(self, iters=20):
        check(lib().sw2d_batch_autotune(self.h_, iters))

    def step(self, n):
        check(lib().sw2d_batch_step(self.h_, int(n)))

    def sync(self):
        check(lib().sw2d_batch_sync(self.h_))

    def idle(self):
        v = C.c_int()
        check(lib().sw2d_batch_idle(self.h_, C.byref(v)))
        return bool(v.value)

    def status(self):
        n = len(self.shapes)
        done = np.zeros(max(n, 1), dtype=np.int32)
        it = np.zeros(max(n, 1), dtype=np.int64)
        nn = C.c_int()
        check(lib().sw2d_batch_status(self.h_, done.ctypes.data_as(C.POINTER(C.c_int)),
                                      it.ctypes.data_as(C.POINTER(C.c_int64)), max(n, 1),
                                      C.byref(nn)))
        return done[:nn.value], it[:nn.value]


class Server:
    """Serving loop with continuous batching (sw2d_server_* in include/sw2d.h)."""

    def __init__(self, precision=FP32, mem_budget=1 << 30, max_slots=0, active_cells=0):
        h = C.c_void_p()
        check(lib().sw2d_server_create(precision, int(mem_budget), int(max_slots), C.byref(h)))
        self.h_ = h
        self.precision = precision
        self.shapes = []
        if active_cells:
            self.set_active_cells(active_cells)

    def set_active_cells(self, cells):
        """Activation target (0: every admitted request active)."""
        check(lib().sw2d_server_set_active_cells(self.h_, int(cells)))

    def close(self):
        if getattr(self, "h_", None):
            lib().sw2d_server_destroy(self.h_)
            self.h_ = None

    def __del__(self):
        self.close()

    def submit(self, z, h0, dx, *, rain=None, dtrain=300.0, dy=None, nodata=-9999.0, n_cell=None,
               **params):
        """Queue a request; returns its id.  rain: table in mm per dtrain."""
        z = np.ascontiguousarray(z, dtype=np.float64)
        h0 = np.ascontiguousarray(h0, dtype=np.float64)
        g = Grid(z.shape[0], z.shape[1], float(dx), float(dy if dy is not None else dx),
                 float(nodata), _dp(z), _dp(h0), None)
        if n_cell is not None:
            n_cell = np.ascontiguousarray(n_cell, dtype=np.float64)
            g.n_cell = _dp(n_cell)
        params.setdefault("precision", self.precision)
        p = default_params(**params)
        r = None if rain is None else np.ascontiguousarray(rain, dtype=np.float64)
        out = C.c_int()
        check(lib().sw2d_server_submit(self.h_, C.byref(g), C.byref(p),
                                       None if r is None else _dp(r), 0 if r is None else len(r),
                                       float(dtrain), C.byref(out)))
        self.shapes.append(z.shape)
        return out.value

    def advance(self, n_iter):
        """Retire, activate, step n_iter iterations, admit; returns (admitted
        and not finished, queued for admission)."""
        nr, nq = C.c_int(), C.c_int()
        check(lib().sw2d_server_advance(self.h_, int(n_iter), C.byref(nr), C.byref(nq)))
        return nr.value, nq.value

    def info(self, rid):
        i = RequestInfo()
        check(lib().sw2d_server_info(self.h_, int(rid), C.byref(i)))
        return {k: getattr(i, k) for k, _ in RequestInfo._fields_}

    def read(self, rid, field):
        r, c = self.shapes[rid]
        shape = {F_H: (r, c), F_U: (r + 1, c), F_V: (r, c + 1)}[field]
        out = np.empty(shape, dtype=np.float64)
        check(lib().sw2d_server_read(self.h_, int(rid), field, _dp(out)))
        return out

    def stats(self, rid):
        st = Stats()
        check(lib().sw2d_server_stats(self.h_, int(rid), C.byref(st)))
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def release(self, rid):
        check(lib().sw2d_server_release(self.h_, int(rid)))
