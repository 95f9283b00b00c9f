"""ctypes binding of libsbp_hip.so (the C-ABI of include/sbp_hip.h).

Mirrors the reference's time-stepping interface for this path:

* ``rhs(ctx, t, v_src, v_dst)`` fills ``v_dst`` with F(t, v_src) in place, like
  the ``PetscErrorCode rhs(TS, PetscReal t, Vec v_src, Vec v_dst, void *ctx)``
  callbacks (e.g. demo/wave/wave_eq_sim.cpp:291-308);
* ``ts_rk4(ctx, t_end, dt, v)`` advances ``v`` in place to ``t_end`` with
  classical RK4 under PETSc's TS_EXACTFINALTIME_MATCHSTEP step schedule, as
  ts_rk4 does (src/time_stepping/ts_rk.cpp:15-39,54-65);
* ``rk4_custom(ctx, t_end, dt, v)`` is the reference's own RK4_custom loop
  (include/time_stepping/time_stepping.h:98-152).

Vectors are float64 numpy arrays in the reference AoS layout v[j][i][c].
The library must be built (``build.py``); nothing here falls back to CPU code.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import problems as P

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsbp_hip.so")
CORR_INPUTS, CORR_OUTPUTS = 8, 3  # SBP_CORR_INPUTS, SBP_CORR_OUTPUTS of sbp_hip.h

_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)


class SbpProblem(C.Structure):
    _fields_ = [("physics", C.c_int), ("order", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("xl", C.c_double), ("xr", C.c_double), ("yl", C.c_double), ("yr", C.c_double),
                ("ax", C.c_double), ("ay", C.c_double),
                ("material", C.c_int), ("bc", C.c_int), ("precision", C.c_int),
                ("rho_inv", _DP), ("bulk", _DP), ("vel_x", _DP), ("vel_y", _DP)]


_ARRAY_FIELDS = ("rho_inv", "bulk", "vel_x", "vel_y")  # copied by sbp_create / sbp_batch_create


class SbpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sbp error %d: %s" % (code, msg))
        self.code = code


EXPORTS = {
    # name: (restype, argtypes)
    "sbp_create": (C.c_int, [C.POINTER(SbpProblem), C.POINTER(C.c_void_p)]),
    "sbp_destroy": (C.c_int, [C.c_void_p]),
    "sbp_components": (C.c_int, [C.c_void_p]),
    "sbp_vector_length": (C.c_long, [C.c_void_p]),
    "sbp_set_tile_rows": (C.c_int, [C.c_void_p, C.c_int]),
    "sbp_set_fused": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "sbp_get_fused": (C.c_int, [C.c_void_p]),
    "sbp_set_store_policy": (C.c_int, [C.c_void_p, C.c_int]),
    "sbp_tune_fused": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_int)]),
    "sbp_batch_set_store_policy": (C.c_int, [C.c_void_p, C.c_int]),
    "sbp_rhs": (C.c_int, [C.c_void_p, C.c_double, _DP, _DP]),
    "sbp_rk4_advance": (C.c_int, [C.c_void_p, _DP, C.c_double, C.c_double, C.c_long]),
    "sbp_set_state": (C.c_int, [C.c_void_p, _DP]),
    "sbp_get_state": (C.c_int, [C.c_void_p, _DP]),
    "sbp_step": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_long]),
    "sbp_rk45": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_long,
                           C.POINTER(C.c_long), C.POINTER(C.c_long), _DP]),
    "sbp_ts_rk4": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_long), _DP]),
    "sbp_matchstep_schedule": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, _DP, C.POINTER(C.c_long),
                                         C.POINTER(C.c_int), _DP]),
    "sbp_rk4_custom": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_long), _DP, _DP]),
    "sbp_adapt_basic": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int), _DP]),
    "sbp_synchronize": (C.c_int, [C.c_void_p]),
    "sbp_time_steps": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_long, _DP]),
    "sbp_time_stages": (C.c_int, [C.c_void_p, C.c_double, C.c_long, _DP]),
    "sbp_energy": (C.c_int, [C.c_void_p, _DP, _DP]),
    "sbp_state_energy": (C.c_int, [C.c_void_p, _DP]),
    "sbp_hnorm": (C.c_int, [C.c_void_p, _DP, _DP]),
    "sbp_error_norms": (C.c_int, [C.c_void_p, _DP, _DP, _DP, _DP]),
    "sbp_error_norms_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_double, C.c_double,
                                          C.c_void_p, C.c_void_p]),
    "sbp_batch_create": (C.c_int, [C.POINTER(SbpProblem), C.c_int, C.POINTER(C.c_void_p)]),
    "sbp_batch_destroy": (C.c_int, [C.c_void_p]),
    "sbp_batch_count": (C.c_int, [C.c_void_p]),
    "sbp_batch_set_state": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "sbp_batch_get_state": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "sbp_batch_step": (C.c_int, [C.c_void_p, _DP, C.c_long]),
    "sbp_batch_synchronize": (C.c_int, [C.c_void_p]),
    "sbp_batch_time_steps": (C.c_int, [C.c_void_p, _DP, C.c_long, _DP]),
    "sbp_batch_energy": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "sbp_batch_num_configs": (C.c_int, []),
    "sbp_batch_set_config": (C.c_int, [C.c_void_p, C.c_int]),
    "sbp_batch_get_config": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sbp_batch_autotune": (C.c_int, [C.c_void_p, _DP, C.POINTER(C.c_int)]),
    "sbp_get_state_strided": (C.c_int, [C.c_void_p, C.c_int, _DP]),
    "sbp_batch_get_state_strided": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _DP]),
    "sbp_corr_param_count": (C.c_int, []),
    "sbp_corr_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "sbp_corr_destroy": (C.c_int, [C.c_void_p]),
    "sbp_corr_set_params": (C.c_int, [C.c_void_p, _FP, C.c_int]),
    "sbp_corr_get_params": (C.c_int, [C.c_void_p, _FP, C.c_int]),
    "sbp_corr_set_normalization": (C.c_int, [C.c_void_p, _FP, _FP, _FP]),
    "sbp_corr_set_band": (C.c_int, [C.c_void_p, C.c_int]),
    "sbp_corr_get_band": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "sbp_corr_forward": (C.c_int, [C.c_void_p, _FP, C.c_int, _FP]),
    "sbp_corr_gradients": (C.c_int, [C.c_void_p, _FP, _FP, C.c_int, _FP, _DP]),
    "sbp_corr_load_dataset": (C.c_int, [C.c_void_p, _FP, _FP, C.c_int]),
    "sbp_corr_train_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_float, _DP]),
    "sbp_corr_synchronize": (C.c_int, [C.c_void_p]),
    "sbp_corr_advance": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_long, C.c_double,
                                   C.POINTER(C.c_int)]),
    "sbp_corr_advance_batch": (C.c_int, [C.c_void_p, C.c_void_p, _DP, C.c_long, C.c_double,
                                         C.POINTER(C.c_int)]),
    "sbp_last_error": (C.c_char_p, []),
    "sbp_device_arch": (C.c_char_p, []),
    "sbp_version": (C.c_int, []),
    "sbp_copy_bandwidth": (C.c_int, [C.c_long, C.c_int, _DP]),
    "sbp_stream_bandwidth": (C.c_int, [C.c_long, C.c_int, C.c_int, C.c_int, _DP]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libsbp_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libsbp_hip.so is not built (run __graft_entry__.build()): " + path)
    lib = C.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise SbpError(rc, load_library().sbp_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def _vec(v, n):
    a = np.ascontiguousarray(v, dtype=np.float64)
    if a.size != n:
        raise ValueError("vector has %d values, the problem needs %d" % (a.size, n))
    return a.reshape(-1)


class Context:
    """One problem on the GPU (the reference's AppCtx + DMDA, sbp_ctx in C)."""

    def __init__(self, spec: P.ProblemSpec):
        lib = load_library()
        self.spec = spec
        self._keep = []
        s = _problem_struct(spec, self._keep)
        h = C.c_void_p()
        _check(lib.sbp_create(C.byref(s), C.byref(h)))
        self._h = h
        self._keep = []  # arrays are copied by sbp_create
        self.n = lib.sbp_vector_length(h)

    # -- lifecycle -------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            load_library().sbp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- reference-shaped calls -------------------------------------------
    def rhs(self, t: float, v: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        src = _vec(v, self.n)
        dst = np.empty(self.n) if out is None else out
        _check(load_library().sbp_rhs(self._h, t, _ptr(src), _ptr(dst)))
        return dst

    def rk4_advance(self, v: np.ndarray, t0: float, dt: float, nsteps: int) -> np.ndarray:
        out = np.array(_vec(v, self.n), copy=True)
        _check(load_library().sbp_rk4_advance(self._h, _ptr(out), t0, dt, int(nsteps)))
        return out

    # -- resident state ---------------------------------------------------
    def set_state(self, v: np.ndarray):
        _check(load_library().sbp_set_state(self._h, _ptr(_vec(v, self.n))))

    def get_state(self) -> np.ndarray:
        out = np.empty(self.n)
        _check(load_library().sbp_get_state(self._h, _ptr(out)))
        return out

    def step(self, t0: float, dt: float, nsteps: int):
        _check(load_library().sbp_step(self._h, t0, dt, int(nsteps)))

    def synchronize(self):
        _check(load_library().sbp_synchronize(self._h))

    def ts_rk4(self, t0: float, t_end: float, dt: float):
        """TSRK4 + MATCHSTEP of the resident state; returns (steps, final time)."""
        n, tf = C.c_long(), C.c_double()
        _check(load_library().sbp_ts_rk4(self._h, t0, t_end, dt, C.byref(n), C.byref(tf)))
        return n.value, tf.value

    def rk4_custom(self, t_end: float, dt: float):
        """RK4_custom's loop on the resident state; returns (tlen, dt used, final t)."""
        n, h, tf = C.c_long(), C.c_double(), C.c_double()
        _check(load_library().sbp_rk4_custom(self._h, t_end, dt, C.byref(n), C.byref(h), C.byref(tf)))
        return n.value, h.value, tf.value

    def rk45(self, t0: float, t_end: float, dt0: float, atol: float = 1e-4, rtol: float = 1e-4,
             max_steps: int = 1000000):
        """Adaptive RKF45 of the resident state to t_end; (accepted, rejected, dt_next)."""
        a, r, d = C.c_long(), C.c_long(), C.c_double()
        _check(load_library().sbp_rk45(self._h, t0, t_end, dt0, atol, rtol, int(max_steps), C.byref(a),
                                       C.byref(r), C.byref(d)))
        return a.value, r.value, d.value

    def time_steps(self, t0: float, dt: float, nsteps: int) -> float:
        ms = C.c_double()
        _check(load_library().sbp_time_steps(self._h, t0, dt, int(nsteps), C.byref(ms)))
        return ms.value

    def time_stages(self, dt: float, nsteps: int) -> np.ndarray:
        out = np.zeros(4)
        _check(load_library().sbp_time_stages(self._h, dt, int(nsteps), _ptr(out)))
        return out

    def get_state_strided(self, stride: int) -> np.ndarray:
        """State at every stride-th point (injection onto the nested grid)."""
        cx = (self.spec.nx - 1) // stride + 1
        cy = (self.spec.ny - 1) // stride + 1
        out = np.empty(cx * cy * self.spec.components)
        _check(load_library().sbp_get_state_strided(self._h, int(stride), _ptr(out)))
        return out

    def set_tile_rows(self, rows: int):
        _check(load_library().sbp_set_tile_rows(self._h, int(rows)))

    def set_fused(self, on: bool = True, rows: int = 0):
        """Two-stage stepping: stages 1-2 and 3-4 in one launch each, interior
        tiles and boundary-frame patches alike (2D acoustics with any material
        model, 2D advection); rows = tile height, 0 = chosen from the grid."""
        _check(load_library().sbp_set_fused(self._h, int(bool(on)), int(rows)))

    @property
    def fused(self) -> bool:
        return bool(load_library().sbp_get_fused(self._h))

    def set_store_policy(self, policy: int):
        """Two-stage tile stores: 0 by streamed size (default), 1 cached, 2 nontemporal."""
        _check(load_library().sbp_set_store_policy(self._h, int(policy)))

    def tune_fused(self, dt: float) -> int:
        """Time the two-stage tile heights on this grid and keep the fastest
        (state untouched); returns the rows chosen (0: one launch per stage)."""
        r = C.c_int()
        _check(load_library().sbp_tune_fused(self._h, float(dt), C.byref(r)))
        return r.value

    # -- norms -------------------------------------------------------------
    def energy(self, v: np.ndarray) -> float:
        e = C.c_double()
        _check(load_library().sbp_energy(self._h, _ptr(_vec(v, self.n)), C.byref(e)))
        return e.value

    def state_energy(self) -> float:
        e = C.c_double()
        _check(load_library().sbp_state_energy(self._h, C.byref(e)))
        return e.value

    def hnorm(self, v: np.ndarray) -> float:
        e = C.c_double()
        _check(load_library().sbp_hnorm(self._h, _ptr(_vec(v, self.n)), C.byref(e)))
        return e.value

    def error_norms(self, v: np.ndarray, w: np.ndarray):
        l2, mx = C.c_double(), C.c_double()
        _check(load_library().sbp_error_norms(self._h, _ptr(_vec(v, self.n)), _ptr(_vec(w, self.n)),
                                              C.byref(l2), C.byref(mx)))
        return l2.value, mx.value


def _problem_struct(spec, keep):
    s = SbpProblem()
    for f, _ in SbpProblem._fields_:
        if f in _ARRAY_FIELDS:
            arr = getattr(spec, f, None)
            if arr is None:
                setattr(s, f, None)
            else:
                a = np.ascontiguousarray(arr, dtype=np.float64)
                keep.append(a)
                setattr(s, f, _ptr(a))
        else:
            setattr(s, f, getattr(spec, f))
    return s


class Batch:
    """Ensemble of independent 2D acoustic problems advanced together
    (sbp_batch in C): one persistent kernel per RK4 stage, graph-replayed steps."""

    def __init__(self, specs):
        lib = load_library()
        self.specs = list(specs)
        keep = []
        arr = (SbpProblem * len(self.specs))()
        for k, sp in enumerate(self.specs):
            arr[k] = _problem_struct(sp, keep)
        h = C.c_void_p()
        _check(lib.sbp_batch_create(arr, len(self.specs), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load_library().sbp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return len(self.specs)

    def _index(self, k):
        if not 0 <= k < len(self.specs):
            raise SbpError(-1, "instance index out of range")
        return k

    def set_state(self, k: int, v: np.ndarray):
        n = self.specs[self._index(k)].length
        _check(load_library().sbp_batch_set_state(self._h, k, _ptr(_vec(v, n))))

    def get_state(self, k: int) -> np.ndarray:
        out = np.empty(self.specs[self._index(k)].length)
        _check(load_library().sbp_batch_get_state(self._h, k, _ptr(out)))
        return out

    def get_state_strided(self, k: int, stride: int) -> np.ndarray:
        sp = self.specs[self._index(k)]
        cx = (sp.nx - 1) // stride + 1
        cy = (sp.ny - 1) // stride + 1
        out = np.empty(cx * cy * 3)
        _check(load_library().sbp_batch_get_state_strided(self._h, k, int(stride), _ptr(out)))
        return out

    def _dt(self, dt):
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (len(self.specs),)))
        return d

    def step(self, dt, nsteps: int):
        d = self._dt(dt)
        _check(load_library().sbp_batch_step(self._h, _ptr(d), int(nsteps)))

    def synchronize(self):
        _check(load_library().sbp_batch_synchronize(self._h))

    def time_steps(self, dt, nsteps: int) -> float:
        d = self._dt(dt)
        ms = C.c_double()
        _check(load_library().sbp_batch_time_steps(self._h, _ptr(d), int(nsteps), C.byref(ms)))
        return ms.value

    def energy(self, k: int) -> float:
        e = C.c_double()
        _check(load_library().sbp_batch_energy(self._h, k, C.byref(e)))
        return e.value

    @staticmethod
    def num_configs() -> int:
        return load_library().sbp_batch_num_configs()

    def set_config(self, cfg: int):
        _check(load_library().sbp_batch_set_config(self._h, int(cfg)))

    def get_config(self):
        v = [C.c_int() for _ in range(5)]
        _check(load_library().sbp_batch_get_config(self._h, *[C.byref(x) for x in v]))
        return dict(config=v[0].value, tile_width=v[1].value, tile_rows=v[2].value,
                    blocks_per_cu=v[3].value, two_stage=bool(v[4].value))

    def set_store_policy(self, policy: int):
        """Two-stage tile stores: 0 by streamed size (default), 1 cached, 2 nontemporal."""
        _check(load_library().sbp_batch_set_store_policy(self._h, int(policy)))

    def autotune(self, dt) -> int:
        d = self._dt(dt)
        c = C.c_int()
        _check(load_library().sbp_batch_autotune(self._h, _ptr(d), C.byref(c)))
        return c.value


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_FP)


def _f32(a, shape=None):
    x = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and x.shape != tuple(shape):
        raise ValueError("array has shape %s, expected %s" % (x.shape, tuple(shape)))
    return x


class Corrector:
    """Learned coarse-grid corrector (sbp_corrector in C): a 3-layer 3x3
    conv net on bf16 MFMA; raw inputs (u, v, p, rho, K), output the
    correction of (u, v, p) in units of out_scale."""

    def __init__(self, nx: int, ny: int, batch: int = 1):
        lib = load_library()
        self.nx, self.ny, self.batch = nx, ny, batch
        h = C.c_void_p()
        _check(lib.sbp_corr_create(nx, ny, batch, C.byref(h)))
        self._h = h
        self.n_params = lib.sbp_corr_param_count()

    def close(self):
        if getattr(self, "_h", None):
            load_library().sbp_corr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        _check(load_library().sbp_corr_get_params(self._h, _fptr(out), self.n_params))
        return out

    @params.setter
    def params(self, p):
        x = _f32(p, (self.n_params,))
        _check(load_library().sbp_corr_set_params(self._h, _fptr(x), self.n_params))

    def set_normalization(self, in_shift, in_scale, out_scale):
        a, b, c = _f32(in_shift, (CORR_INPUTS,)), _f32(in_scale, (CORR_INPUTS,)), _f32(out_scale, (3,))
        _check(load_library().sbp_corr_set_normalization(self._h, _fptr(a), _fptr(b), _fptr(c)))

    @property
    def band(self) -> int:
        """Points nearer a wall than this keep the plain scheme's values."""
        v = C.c_int()
        _check(load_library().sbp_corr_get_band(self._h, C.byref(v)))
        return v.value

    @band.setter
    def band(self, band: int):
        _check(load_library().sbp_corr_set_band(self._h, int(band)))

    def forward(self, raw) -> np.ndarray:
        raw = _f32(raw)
        nb = raw.shape[0]
        _f32(raw, (nb, self.ny, self.nx, CORR_INPUTS))
        out = np.empty((nb, self.ny, self.nx, 3), dtype=np.float32)
        _check(load_library().sbp_corr_forward(self._h, _fptr(raw), nb, _fptr(out)))
        return out

    def gradients(self, raw, target):
        raw = _f32(raw)
        nb = raw.shape[0]
        _f32(raw, (nb, self.ny, self.nx, CORR_INPUTS))
        tgt = _f32(target, (nb, self.ny, self.nx, 3))
        g = np.empty(self.n_params, dtype=np.float32)
        loss = C.c_double()
        _check(load_library().sbp_corr_gradients(self._h, _fptr(raw), _fptr(tgt), nb, _fptr(g), C.byref(loss)))
        return loss.value, g

    def load_dataset(self, raw, target):
        raw = _f32(raw)
        n = raw.shape[0]
        _f32(raw, (n, self.ny, self.nx, CORR_INPUTS))
        tgt = _f32(target, (n, self.ny, self.nx, 3))
        _check(load_library().sbp_corr_load_dataset(self._h, _fptr(raw), _fptr(tgt), n))

    def train_step(self, idx, lr: float) -> float:
        ix = np.ascontiguousarray(idx, dtype=np.int32)
        loss = C.c_double()
        _check(load_library().sbp_corr_train_step(self._h, ix.ctypes.data_as(C.POINTER(C.c_int)), len(ix),
                                                   float(lr), C.byref(loss)))
        return loss.value

    def advance(self, ctx: "Context", t0: float, dt: float, nsteps: int, bound: float) -> int:
        rej = C.c_int()
        _check(load_library().sbp_corr_advance(self._h, ctx._h, t0, dt, int(nsteps), float(bound), C.byref(rej)))
        return rej.value

    def advance_batch(self, batch: "Batch", dt, nsteps: int, bound: float) -> np.ndarray:
        d = batch._dt(dt)
        rej = np.zeros(len(batch), dtype=np.int32)
        _check(load_library().sbp_corr_advance_batch(self._h, batch._h, _ptr(d), int(nsteps), float(bound),
                                                     rej.ctypes.data_as(C.POINTER(C.c_int))))
        return rej


def rhs(ctx: Context, t: float, v_src: np.ndarray, v_dst: np.ndarray) -> int:
    """PETSc-callback shaped right-hand side: writes F(t, v_src) into v_dst."""
    if v_dst.dtype != np.float64 or not v_dst.flags.c_contiguous:
        raise ValueError("v_dst must be a contiguous float64 array")
    ctx.rhs(t, v_src, out=v_dst.reshape(-1))
    return 0


def ts_rk4(ctx: Context, t_end: float, dt: float, v: np.ndarray, t0: float = 0.0) -> float:
    """Advance v in place from t0 to t_end with TSRK4 under
    TS_EXACTFINALTIME_MATCHSTEP (PETSc's step schedule, sbp_ts_rk4); returns
    the final time."""
    ctx.set_state(v)
    _, tf = ctx.ts_rk4(t0, t_end, dt)
    v.reshape(-1)[:] = ctx.get_state()
    return tf


def rk4_custom(ctx: Context, t_end: float, dt: float, v: np.ndarray, verbose: bool = True):
    """RK4_custom(da, Tend, dt, v, rhs, ctx) (time_stepping.h:98-152): v advanced
    in place over round(Tend/dt) steps from t = 0, with the reference's prints;
    returns (tlen, dt used, final t)."""
    tlen = int(round(t_end / dt))
    if verbose and abs(tlen * dt - t_end) > 1e-14:
        print("Warning: Non-matching time step. dt*tlen = %.12f.\nChanging timestep from %e to %e."
              % (dt * tlen, dt, t_end / tlen))
    ctx.set_state(v)
    n, h, tf = ctx.rk4_custom(t_end, dt)
    if verbose:
        print("tlen: %d, dt: %e" % (n, h))
        print("Final t: %.8e" % tf)
    v.reshape(-1)[:] = ctx.get_state()
    return n, h, tf


def matchstep_schedule(t_end: float, dt: float, t0: float = 0.0, max_runs: int = 64):
    """ts_rk4's step sizes from t0 to t_end (host only): [(dt, count), ...]."""
    rd = np.empty(max_runs)
    rl = np.empty(max_runs, dtype=np.int64)
    nr, tf = C.c_int(), C.c_double()
    _check(load_library().sbp_matchstep_schedule(t0, t_end, dt, max_runs, _ptr(rd),
                                                 rl.ctypes.data_as(C.POINTER(C.c_long)), C.byref(nr), C.byref(tf)))
    return [(float(rd[k]), int(rl[k])) for k in range(nr.value)]


def adapt_basic(enorm: float, h: float, order: int = 5, prev_rejected: bool = False):
    """The library's TSAdaptChoose_Basic decision (host only): (accept, next_h)."""
    a, nh = C.c_int(), C.c_double()
    _check(load_library().sbp_adapt_basic(enorm, h, order, int(prev_rejected), C.byref(a), C.byref(nh)))
    return bool(a.value), nh.value


def ts_rk45(ctx: Context, t_end: float, dt: float, v: np.ndarray, t0: float = 0.0, atol: float = 1e-4,
            rtol: float = 1e-4):
    """ts_rk45 (ts_rk.cpp:41-52): adaptive RKF45 of v in place from t0 to
    t_end, first step dt; returns (accepted, rejected)."""
    ctx.set_state(v)
    a, r, _ = ctx.rk45(t0, t_end, dt, atol, rtol)
    v.reshape(-1)[:] = ctx.get_state()
    return a, r


def error_norms_noalloc(v_ptr: int, w_ptr: int, n: int, hx: float, hy: float, out_ptr: int,
                        stream: int = 0):
    """sbp_error_norms_noalloc on device pointers (integers): out[0] = l2,
    out[1] = max, written asynchronously on `stream` without any allocation."""
    _check(load_library().sbp_error_norms_noalloc(C.c_void_p(v_ptr), C.c_void_p(w_ptr), int(n), float(hx),
                                                  float(hy), C.c_void_p(out_ptr), C.c_void_p(stream or None)))


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 5) -> float:
    """GB/s of a plain device copy of nbytes (read + write counted)."""
    g = C.c_double()
    _check(load_library().sbp_copy_bandwidth(int(nbytes), int(reps), C.byref(g)))
    return g.value


def stream_bandwidth(plane_bytes: int, n_read: int, n_write: int, reps: int = 5) -> float:
    """GB/s (read + write counted) of a flat elementwise kernel reading n_read
    and writing n_write planes of plane_bytes each."""
    g = C.c_double()
    _check(load_library().sbp_stream_bandwidth(int(plane_bytes), int(n_read), int(n_write), int(reps), C.byref(g)))
    return g.value


def device_arch() -> str:
    return load_library().sbp_device_arch().decode()
