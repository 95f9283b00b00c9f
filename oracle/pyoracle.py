"""ctypes binding of oracle/liborc.so and oracle/libcpubase.so (test infrastructure).

Problem descriptions are duck-typed: any object with the fields of
``sbp_problem`` (physics, order, nx, ny, xl, xr, yl, yr, ax, ay, material, bc,
precision, rho_inv, bulk, vel_x, vel_y) is accepted; missing array fields are NULL.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class OrcProblem(C.Structure):
    _fields_ = [("physics", C.c_int), ("order", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("xl", C.c_double), ("xr", C.c_double), ("yl", C.c_double), ("yr", C.c_double),
                ("ax", C.c_double), ("ay", C.c_double),
                ("material", C.c_int), ("bc", C.c_int), ("precision", C.c_int),
                ("rho_inv", C.POINTER(C.c_double)), ("bulk", C.POINTER(C.c_double)),
                ("vel_x", C.POINTER(C.c_double)), ("vel_y", C.POINTER(C.c_double))]

_ARRAY_FIELDS = ("rho_inv", "bulk", "vel_x", "vel_y")


class OrcOps(C.Structure):
    _fields_ = [("order", C.c_int), ("r", C.c_int), ("cls", C.c_int), ("w", C.c_int),
                ("inner", C.c_double * 7), ("Q", (C.c_double * 9) * 6),
                ("hq", C.c_double * 6), ("hqinv", C.c_double * 6)]


_libs = {}
_DP = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(name):
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        P = C.POINTER(OrcProblem)
        lib.orc_get_ops.argtypes = [C.c_int, C.POINTER(OrcOps)]
        lib.orc_validate.argtypes = [P]
        lib.orc_d1.argtypes = [C.POINTER(OrcOps), _DP, C.c_long, C.c_int, C.c_int, C.c_double]
        lib.orc_d1.restype = C.c_double
        lib.orc_rhs.argtypes = [P, C.c_double, _DP, _DP]
        lib.orc_rk4.argtypes = [P, _DP, C.c_double, C.c_double, C.c_long]
        lib.orc_rk4_par.argtypes = [P, _DP, C.c_double, C.c_double, C.c_long, C.c_int]
        lib.orc_energy.argtypes = [P, _DP]
        lib.orc_energy.restype = C.c_double
        lib.orc_hnorm.argtypes = [P, _DP]
        lib.orc_hnorm.restype = C.c_double
        lib.orc_error_norms.argtypes = [P, _DP, _DP, _DP, _DP]
        _LP = C.POINTER(C.c_long)
        lib.orc_rk45.argtypes = [P, _DP, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_long,
                                 _LP, _LP, _DP]
        lib.orc_rkf45_tableau.argtypes = [_DP, _DP, _DP, _DP]
        lib.orc_ts_rk4.argtypes = [P, _DP, C.c_double, C.c_double, C.c_double, _LP, _DP, C.c_int]
        lib.orc_matchstep_steps.argtypes = [C.c_double, C.c_double, C.c_double, C.c_long, _DP, _DP]
        lib.orc_adapt_basic.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), _DP]
        lib.orc_adapt_basic.restype = None
        _libs[name] = lib
    return _libs[name]


def lib(fast: bool = False):
    return _load("libcpubase.so" if fast else "liborc.so")


class _Desc:
    """Keeps the material arrays alive while the struct points at them."""

    def __init__(self, spec):
        self.keep = []
        s = OrcProblem()
        for f, _ in OrcProblem._fields_:
            if f in _ARRAY_FIELDS:
                arr = getattr(spec, f, None)
                if arr is None:
                    setattr(s, f, None)
                else:
                    a = np.ascontiguousarray(arr, dtype=np.float64)
                    self.keep.append(a)
                    setattr(s, f, a.ctypes.data_as(_DP))
            else:
                setattr(s, f, getattr(spec, f))
        self.s = s

    @property
    def ref(self):
        return C.byref(self.s)


def _ptr(a):
    return a.ctypes.data_as(_DP)


def ops(order: int) -> OrcOps:
    o = OrcOps()
    if lib().orc_get_ops(order, C.byref(o)):
        raise ValueError(order)
    return o


def validate(spec) -> bool:
    return lib().orc_validate(_Desc(spec).ref) == 0


def d1(order: int, v: np.ndarray, i: int, hi: float = 1.0) -> float:
    o = ops(order)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_d1(C.byref(o), _ptr(v), 1, len(v), i, hi)


def rhs(spec, t: float, v: np.ndarray) -> np.ndarray:
    d = _Desc(spec)
    src = np.ascontiguousarray(v, dtype=np.float64)
    dst = np.zeros_like(src)
    if lib().orc_rhs(d.ref, t, _ptr(src), _ptr(dst)):
        raise ValueError("orc_rhs rejected the problem")
    return dst


def rk4(spec, v: np.ndarray, t0: float, dt: float, nsteps: int, threads: int = 1,
        fast: bool = False) -> np.ndarray:
    d = _Desc(spec)
    out = np.array(v, dtype=np.float64, copy=True)
    rc = lib(fast).orc_rk4_par(d.ref, _ptr(out), t0, dt, nsteps, threads)
    if rc:
        raise ValueError("orc_rk4 rejected the problem")
    return out


def ts_rk4(spec, v: np.ndarray, t_end: float, dt: float, t0: float = 0.0, threads: int = 1):
    """ts_rk4 (ts_rk.cpp:54-65): RK4 under TS_EXACTFINALTIME_MATCHSTEP from t0
    to t_end, first step dt; returns (v, steps taken, final time)."""
    d = _Desc(spec)
    out = np.array(v, dtype=np.float64, copy=True)
    n, tf = C.c_long(), C.c_double()
    if lib().orc_ts_rk4(d.ref, _ptr(out), t0, t_end, dt, C.byref(n), C.byref(tf), threads):
        raise ValueError("orc_ts_rk4 rejected the problem")
    return out, n.value, tf.value


def matchstep_steps(t0: float, t_end: float, dt: float, max_steps: int = 1 << 20):
    """Step sizes of a fixed-step MATCHSTEP run; returns (steps, final time)."""
    buf = np.empty(max_steps)
    tf = C.c_double()
    n = lib().orc_matchstep_steps(t0, t_end, dt, max_steps, _ptr(buf), C.byref(tf))
    if n < 0:
        raise ValueError("more than max_steps steps")
    return buf[:n].copy(), tf.value


def adapt_basic(enorm: float, h: float, order: int = 5, prev_rejected: bool = False):
    """TSAdaptChoose_Basic's (accept, next_h) for one attempt."""
    acc, nh = C.c_int(), C.c_double()
    lib().orc_adapt_basic(enorm, h, order, int(prev_rejected), C.byref(acc), C.byref(nh))
    return bool(acc.value), nh.value


def rk45(spec, v: np.ndarray, t0: float, t_end: float, dt0: float, atol: float = 1e-4,
         rtol: float = 1e-4, max_steps: int = 1000000, check: bool = True):
    """Adaptive RKF45 (orc_rk45); returns (v, accepted, rejected, dt_next), and
    the return code as a fifth value when check is False (no exception then:
    -2 = max_steps attempts spent, -3 = more than 10 rejections of one step
    or a non-finite error estimate)."""
    d = _Desc(spec)
    out = np.array(v, dtype=np.float64, copy=True)
    acc, rej, dtn = C.c_long(), C.c_long(), C.c_double()
    rc = lib().orc_rk45(d.ref, _ptr(out), t0, t_end, dt0, atol, rtol, max_steps, C.byref(acc), C.byref(rej),
                        C.byref(dtn))
    if check:
        if rc:
            raise ValueError("orc_rk45 failed (%d)" % rc)
        return out, acc.value, rej.value, dtn.value
    return out, acc.value, rej.value, dtn.value, rc


def rkf45_tableau():
    """(c[6], a[6][5], b5[6], b4[6]) as compiled into the oracle."""
    c, a, b5, b4 = np.zeros(6), np.zeros((6, 5)), np.zeros(6), np.zeros(6)
    lib().orc_rkf45_tableau(_ptr(c), _ptr(a), _ptr(b5), _ptr(b4))
    return c, a, b5, b4


def energy(spec, v: np.ndarray) -> float:
    d = _Desc(spec)
    return lib().orc_energy(d.ref, _ptr(np.ascontiguousarray(v, dtype=np.float64)))


def hnorm(spec, v: np.ndarray) -> float:
    d = _Desc(spec)
    return lib().orc_hnorm(d.ref, _ptr(np.ascontiguousarray(v, dtype=np.float64)))


def error_norms(spec, v: np.ndarray, w: np.ndarray):
    d = _Desc(spec)
    l2 = C.c_double()
    mx = C.c_double()
    lib().orc_error_norms(d.ref, _ptr(np.ascontiguousarray(v, dtype=np.float64)),
                          _ptr(np.ascontiguousarray(w, dtype=np.float64)),
                          C.byref(l2), C.byref(mx))
    return l2.value, mx.value
