"""ctypes binding of the C oracle (liboracle_sw2d.so).

TEST INFRASTRUCTURE ONLY: used by the tests, the smoke check and the CPU
baseline leg of bench.py, never by the product library.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_sw2d.so")

EDGE_NORTH, EDGE_SOUTH, EDGE_WEST, EDGE_EAST = 1, 2, 4, 8
H, U, V, UU, VV, HM, HN, UU1, VV1, UUA, VVA = range(11)


class Params(C.Structure):
    _fields_ = [("rows", C.c_int), ("cols", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dt", C.c_double),
                ("g", C.c_double), ("th", C.c_double), ("manning", C.c_double),
                ("INT", C.c_double), ("INF", C.c_double), ("LWL", C.c_double),
                ("EV_WL_min", C.c_double), ("evaporation_on", C.c_int),
                ("cfl", C.c_double), ("edge_open", C.c_int),
                ("forcing_mode", C.c_int), ("t0", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("t", C.c_double), ("iterations", C.c_longlong),
                ("last_D", C.c_double), ("volume", C.c_double),
                ("rain_in", C.c_double), ("source_in", C.c_double),
                ("evap_out", C.c_double), ("outflow", C.c_double),
                ("clamp_gain", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle library not built: run make -C oracle")
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params), dp, C.POINTER(C.c_ubyte), dp, dp]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_rain.argtypes = [C.c_void_p, dp, C.c_int, C.c_double, dp, C.c_int]
        L.orc_set_sources.argtypes = [C.c_void_p, dp, C.c_int, C.c_int, C.c_double,
                                      C.POINTER(C.c_longlong)]
        L.orc_set_evaporation.argtypes = [C.c_void_p, dp, C.c_int]
        L.orc_set_dt_schedule.argtypes = [C.c_void_p, dp, C.c_int]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_state.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_read.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.orc_stream_rate.restype = C.c_double
        L.orc_stream_rate.argtypes = [C.c_double, dp, C.c_int, C.c_double, C.c_int]
        L.orc_pow_fric.restype = C.c_double
        L.orc_pow_fric.argtypes = [C.c_double]
        L.orc_evap_table_entry.restype = C.c_double
        L.orc_evap_table_entry.argtypes = [C.c_double] * 4
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleSim:
    """One FP64 oracle run.  Arrays are (rows, cols) float64."""

    def __init__(self, z, active, h0, dx, dt, manning, *, dy=None, g=9.80, th=1.0e-4,
                 INT=0.0, INF=0.0, LWL=0.0, EV_WL_min=0.0, evaporation_on=0,
                 cfl=0.0, edge_open=0, forcing_mode=0, t0=0.0, n_cell=None):
        L = lib()
        self.rows, self.cols = z.shape
        p = Params(self.rows, self.cols, dx, dy if dy is not None else dx, dt, g, th,
                   manning, INT, INF, LWL, EV_WL_min, evaporation_on, cfl,
                   edge_open, forcing_mode, t0)
        self._z = np.ascontiguousarray(z, dtype=np.float64)
        self._a = np.ascontiguousarray(active, dtype=np.uint8)
        self._h = np.ascontiguousarray(h0, dtype=np.float64)
        ncp = None
        if n_cell is not None:
            self._n = np.ascontiguousarray(n_cell, dtype=np.float64)
            ncp = _dp(self._n)
        self.h_ = L.orc_create(C.byref(p), _dp(self._z),
                               self._a.ctypes.data_as(C.POINTER(C.c_ubyte)),
                               _dp(self._h), ncp)
        if not self.h_:
            raise ValueError("orc_create failed")

    def __del__(self):
        if getattr(self, "h_", None):
            lib().orc_destroy(self.h_)
            self.h_ = None

    def set_rain(self, table_mm, dtrain, block_map=None, block=0):
        t = np.ascontiguousarray(table_mm, dtype=np.float64)
        bm = None
        if block_map is not None:
            self._bm = np.ascontiguousarray(block_map, dtype=np.float64)
            bm = _dp(self._bm)
        lib().orc_set_rain(self.h_, _dp(t), len(t), dtrain, bm, block)

    def set_sources(self, q, dtoq, cell_index):
        q = np.ascontiguousarray(q, dtype=np.float64)
        ci = np.ascontiguousarray(cell_index, dtype=np.int64)
        lib().orc_set_sources(self.h_, _dp(q), q.shape[0], q.shape[1], dtoq,
                              ci.ctypes.data_as(C.POINTER(C.c_longlong)))

    def set_evaporation(self, evapo):
        e = np.ascontiguousarray(evapo, dtype=np.float64)
        lib().orc_set_evaporation(self.h_, _dp(e), len(e))

    def set_dt_schedule(self, D):
        d = np.ascontiguousarray(D, dtype=np.float64)
        lib().orc_set_dt_schedule(self.h_, _dp(d), len(d))

    def set_state(self, h, U, V):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (h, U, V)]
        lib().orc_set_state(self.h_, *[_dp(a) for a in arrs])

    def step(self, n):
        if lib().orc_step(self.h_, int(n)) != 0:
            raise RuntimeError("orc_step failed")

    def read(self, field):
        R, Cc = self.rows, self.cols
        shape = {H: (R, Cc), UU: (R, Cc), VV: (R, Cc), U: (R + 1, Cc), HM: (R + 1, Cc),
                 UU1: (R + 1, Cc), VVA: (R + 1, Cc), V: (R, Cc + 1), HN: (R, Cc + 1),
                 VV1: (R, Cc + 1), UUA: (R, Cc + 1)}[field]
        out = np.empty(shape)
        lib().orc_read(self.h_, field, _dp(out))
        return out

    def stats(self):
        s = Stats()
        lib().orc_get_stats(self.h_, C.byref(s))
        return {k: getattr(s, k) for k, _ in Stats._fields_}


def stream_rate(time, table, dT, mode=0):
    t = np.ascontiguousarray(table, dtype=np.float64)
    return lib().orc_stream_rate(time, _dp(t), len(t), dT, mode)


def evap_entry(albedo, T, Rg, dtrain):
    return lib().orc_evap_table_entry(albedo, T, Rg, dtrain)


def pow_fric(x):
    """hf^1.3333330 as the oracle evaluates it (orc_pow_fric)."""
    return lib().orc_pow_fric(float(x))
