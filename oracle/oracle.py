"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the serial FP64 oracle (oracle/libns_oracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module. It takes
the same ``ns_config`` structure as the device library (built with nsfsi.make_config).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libns_oracle.so")

FIELD_IDS = {"rho": 0, "rho_u": 1, "rho_v": 2, "rho_w": 3, "rho_E": 4, "u": 5, "v": 6, "w": 7, "p": 8, "T": 9,
             "mu": 10, "kappa": 11}

_lib = None


def build():
    """Compiles the oracle with its Makefile (g++ -O3 -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    H = C.c_void_p
    D = C.POINTER(C.c_double)
    L.ns_oracle_create.argtypes = [C.c_void_p, C.POINTER(H)]
    L.ns_oracle_initialize.argtypes = [H]
    L.ns_oracle_march.argtypes = [H, C.c_int]
    L.ns_oracle_run.argtypes = [H]
    L.ns_oracle_state.argtypes = [H, D, C.POINTER(C.c_uint64), D]
    L.ns_oracle_state.restype = None
    L.ns_oracle_geometry.argtypes = [H, D, D]
    L.ns_oracle_geometry.restype = None
    L.ns_oracle_get_field.argtypes = [H, C.c_int, D]
    L.ns_oracle_set_field.argtypes = [H, C.c_int, D]
    L.ns_oracle_derive_all.argtypes = [H]
    L.ns_oracle_slopes.argtypes = [H, D]
    L.ns_oracle_apply_bcs.argtypes = [H, C.c_double]
    L.ns_oracle_get_node_types.argtypes = [H, C.POINTER(C.c_uint8)]
    L.ns_oracle_counts.argtypes = [H, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.ns_oracle_counts.restype = None
    L.ns_oracle_get_ghosts.argtypes = [H, C.POINTER(C.c_int32), D, D, C.POINTER(C.c_int32)]
    L.ns_oracle_get_filter_nodes.argtypes = [H, C.POINTER(C.c_int32)]
    L.ns_oracle_norm_history.argtypes = [H, C.c_int, D, C.c_int64]
    L.ns_oracle_norm_history.restype = C.c_int64
    L.ns_oracle_last_error.argtypes = [H]
    L.ns_oracle_last_error.restype = C.c_char_p
    L.ns_oracle_destroy.argtypes = [H]
    L.ns_oracle_destroy.restype = None
    L.ns_oracle_lu8_solve.argtypes = [D, D, D]
    L.ns_oracle_lu8_solve.restype = None
    L.ns_oracle_uniform01.argtypes = [C.c_uint64, C.c_uint64]
    L.ns_oracle_uniform01.restype = C.c_double
    L.ns_oracle_set_cr_pow.argtypes = [H, C.c_int]
    L.ns_oracle_pow15.argtypes = [C.c_double, C.c_int]
    L.ns_oracle_pow15.restype = C.c_double
    L.ns_oracle_set_time.argtypes = [H, C.c_double, C.c_uint64]
    L.ns_oracle_filter_now.argtypes = [H]
    L.ns_oracle_rho_filter.argtypes = [H]
    L.ns_oracle_norm_of_change.argtypes = [H, C.c_int]
    L.ns_oracle_norm_of_change.restype = C.c_double
    L.ns_oracle_set_dt_schedule.argtypes = [H, D, C.c_int64]
    L.ns_oracle_update_dt.argtypes = [H, C.c_double]
    L.ns_oracle_update_dt.restype = C.c_double
    L.ns_oracle_tl_create.argtypes = [C.c_void_p, C.POINTER(H)]
    L.ns_oracle_tl_initialize.argtypes = [H]
    L.ns_oracle_tl_set_state.argtypes = [H, D, D]
    L.ns_oracle_tl_get_state.argtypes = [H, D, D]
    L.ns_oracle_tl_set_dt_schedule.argtypes = [H, D, C.c_int64]
    L.ns_oracle_tl_set_re_tangent.argtypes = [H, C.c_double]
    L.ns_oracle_tl_march.argtypes = [H, C.c_int]
    L.ns_oracle_tl_slopes.argtypes = [H, D, D]
    L.ns_oracle_tl_state.argtypes = [H, D, C.POINTER(C.c_uint64), D]
    L.ns_oracle_tl_state.restype = None
    L.ns_oracle_tl_last_error.argtypes = [H]
    L.ns_oracle_tl_last_error.restype = C.c_char_p
    L.ns_oracle_tl_destroy.argtypes = [H]
    L.ns_oracle_tl_destroy.restype = None
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"oracle status {status}: {msg}")
        self.status = status


class Oracle:
    """Serial FP64 restatement of the reference time step (see ns_oracle.cpp)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._h = C.c_void_p()
        s = lib().ns_oracle_create(C.byref(cfg), C.byref(self._h))
        if s != 0:
            raise OracleError(s, "create failed")
        self.NI, self.NJ, self.NK = cfg.NI, cfg.NJ, cfg.NK
        self.N = self.NI * self.NJ * self.NK

    def _check(self, s):
        if s != 0:
            raise OracleError(s, lib().ns_oracle_last_error(self._h).decode())

    def initialize(self):
        self._check(lib().ns_oracle_initialize(self._h))

    def march(self, n=1):
        self._check(lib().ns_oracle_march(self._h, int(n)))

    def run(self):
        self._check(lib().ns_oracle_run(self._h))

    def state(self):
        t, dt = C.c_double(), C.c_double()
        lv = C.c_uint64()
        lib().ns_oracle_state(self._h, C.byref(t), C.byref(lv), C.byref(dt))
        return t.value, lv.value, dt.value

    def geometry(self):
        sp, of = np.zeros(3), np.zeros(3)
        lib().ns_oracle_geometry(self._h, _d(sp), _d(of))
        return sp, of

    def field(self, name) -> np.ndarray:
        which = FIELD_IDS[name] if isinstance(name, str) else int(name)
        out = np.empty(self.N)
        self._check(lib().ns_oracle_get_field(self._h, which, _d(out)))
        return out.reshape(self.NK, self.NJ, self.NI)

    def set_field(self, name, values):
        which = FIELD_IDS[name] if isinstance(name, str) else int(name)
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        assert a.size == self.N
        self._check(lib().ns_oracle_set_field(self._h, which, _d(a)))

    def conserved(self) -> np.ndarray:
        return np.stack([self.field(q) for q in range(5)])

    def derive_all(self):
        self._check(lib().ns_oracle_derive_all(self._h))

    def slopes(self) -> np.ndarray:
        out = np.zeros(5 * self.N)
        self._check(lib().ns_oracle_slopes(self._h, _d(out)))
        return out.reshape(5, self.NK, self.NJ, self.NI)

    def apply_bcs(self, t):
        self._check(lib().ns_oracle_apply_bcs(self._h, float(t)))

    def node_types(self) -> np.ndarray:
        out = np.empty(self.N, dtype=np.uint8)
        self._check(lib().ns_oracle_get_node_types(self._h, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out.reshape(self.NK, self.NJ, self.NI)

    def counts(self):
        g, f, n = C.c_int64(), C.c_int64(), C.c_int64()
        lib().ns_oracle_counts(self._h, C.byref(g), C.byref(f), C.byref(n))
        return g.value, f.value, n.value

    def ghosts(self):
        n = self.counts()[0]
        ijk = np.zeros((n, 3), dtype=np.int32)
        bi, ip = np.zeros((n, 3)), np.zeros((n, 3))
        box = np.zeros((n, 3), dtype=np.int32)
        self._check(lib().ns_oracle_get_ghosts(self._h, ijk.ctypes.data_as(C.POINTER(C.c_int32)), _d(bi), _d(ip),
                                               box.ctypes.data_as(C.POINTER(C.c_int32))))
        return ijk, bi, ip, box

    def filter_nodes(self):
        n = self.counts()[1]
        ijk = np.zeros((n, 3), dtype=np.int32)
        if n:
            self._check(lib().ns_oracle_get_filter_nodes(self._h, ijk.ctypes.data_as(C.POINTER(C.c_int32))))
        return ijk

    def norm_history(self, var):
        n = lib().ns_oracle_norm_history(self._h, int(var), None, 0)
        out = np.zeros(max(n, 0))
        if n > 0:
            lib().ns_oracle_norm_history(self._h, int(var), _d(out), n)
        return out

    def set_old_field(self, q, values):
        """Overwrite the "Old" copy of conserved variable q (the buffer the norm of change reads)."""
        self.set_field(12 + int(q), values)

    def old_field(self, q) -> np.ndarray:
        return self.field(12 + int(q))

    def set_correctly_rounded_pow(self, on=True):
        """x^1.5 of the viscosity law correctly rounded (True) or by std::pow, the reference's call (False)."""
        self._check(lib().ns_oracle_set_cr_pow(self._h, 1 if on else 0))

    def set_time(self, t, level):
        self._check(lib().ns_oracle_set_time(self._h, float(t), int(level)))

    def filter_now(self) -> bool:
        return lib().ns_oracle_filter_now(self._h) == 1

    def rho_filter(self) -> bool:
        """The periodic density filter as the step applies it; True when its schedule fired."""
        return lib().ns_oracle_rho_filter(self._h) == 1

    def norm_of_change(self, q) -> float:
        return lib().ns_oracle_norm_of_change(self._h, int(q))

    def set_dt_schedule(self, dts):
        """The step from time level n takes dts[n] (the device's ns_set_dt_schedule)."""
        a = np.ascontiguousarray(dts, dtype=np.float64)
        self._check(lib().ns_oracle_set_dt_schedule(self._h, _d(a), a.size))

    def update_dt(self, tnow) -> float:
        return lib().ns_oracle_update_dt(self._h, float(tnow))

    def close(self):
        if self._h:
            lib().ns_oracle_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TangentOracle:
    """Tangent-linear form of the oracle: the same set-up and step on (value, tangent) pairs, so each value
    equals the plain oracle's bit for bit and each tangent is the derivative of the computed step along
    the tangent given for the state (at the step sizes the run takes; prescribe them to hold them fixed)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._h = C.c_void_p()
        s = lib().ns_oracle_tl_create(C.byref(cfg), C.byref(self._h))
        if s != 0:
            raise OracleError(s, "create failed")
        self.NI, self.NJ, self.NK = cfg.NI, cfg.NJ, cfg.NK
        self.N = self.NI * self.NJ * self.NK

    def _check(self, s):
        if s != 0:
            raise OracleError(s, lib().ns_oracle_tl_last_error(self._h).decode())

    def initialize(self):
        self._check(lib().ns_oracle_tl_initialize(self._h))

    def set_state(self, values, tangents=None):
        """Conserved state (5, NK, NJ, NI) and its tangent; primitives, transport and dt are derived."""
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        assert v.size == 5 * self.N
        t = None if tangents is None else np.ascontiguousarray(tangents, dtype=np.float64).reshape(-1)
        self._check(lib().ns_oracle_tl_set_state(self._h, _d(v), None if t is None else _d(t)))

    def state_pair(self):
        v, t = np.empty(5 * self.N), np.empty(5 * self.N)
        self._check(lib().ns_oracle_tl_get_state(self._h, _d(v), _d(t)))
        return v.reshape(5, self.NK, self.NJ, self.NI), t.reshape(5, self.NK, self.NJ, self.NI)

    def set_dt_schedule(self, dts):
        a = np.ascontiguousarray(dts, dtype=np.float64)
        self._check(lib().ns_oracle_tl_set_dt_schedule(self._h, _d(a), a.size))

    def set_re_tangent(self, dre):
        """Tangent of the Reynolds number (before set_state, which derives the viscosity from it)."""
        self._check(lib().ns_oracle_tl_set_re_tangent(self._h, float(dre)))

    def march(self, n=1):
        self._check(lib().ns_oracle_tl_march(self._h, int(n)))

    def slopes(self):
        v, t = np.zeros(5 * self.N), np.zeros(5 * self.N)
        self._check(lib().ns_oracle_tl_slopes(self._h, _d(v), _d(t)))
        return v.reshape(5, self.NK, self.NJ, self.NI), t.reshape(5, self.NK, self.NJ, self.NI)

    def state(self):
        t, dt = C.c_double(), C.c_double()
        lv = C.c_uint64()
        lib().ns_oracle_tl_state(self._h, C.byref(t), C.byref(lv), C.byref(dt))
        return t.value, lv.value, dt.value

    def close(self):
        if self._h:
            lib().ns_oracle_tl_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lu8_solve(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(8)
    lib().ns_oracle_lu8_solve(_d(A), _d(b), _d(x))
    return x


def pow15(x, correctly_rounded):
    return lib().ns_oracle_pow15(float(x), 1 if correctly_rounded else 0)


def uniform01(seed, counter):
    return lib().ns_oracle_uniform01(seed, counter)
