"""Problem set-ups: the reference demos and the benchmark workloads.

Each builder returns a ``ProblemSpec`` (the Python image of ``sbp_problem`` in
include/sbp_hip.h) plus the initial state as an AoS array ``v[j][i][c]``
flattened C-order, the reference's single-rank layout
(code/include/grids/layout.h:118-127).

Reference demos (constants from the drivers):
  adv_1D      demo/advection/advection_1D_sim.cpp:59-70,171-188
  adv_2D      demo/advection/advection_2D_sim.cpp:61-75,188-209
  wave        demo/wave/wave_eq_sim.cpp:81-91,211-233 (rho = 2 + xy, forced)
  wave_hom    demo/wave_hom/wave_eq_hom_sim.cpp (rho = 1, unforced)
  reflection  demo/reflection/reflection_sim.cpp:56-66,177-228
Benchmark workloads: seeded 2D acoustic problems on the unit square with wall
boundaries, Gaussian pressure pulses and smooth material fields.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

# sbp_physics / sbp_material / sbp_bc / sbp_precision of include/sbp_hip.h
ADVECTION_1D, ADVECTION_2D, ACOUSTIC_2D, REFLECTION_1D = 1, 2, 3, 4
MAT_UNIT, MAT_REF_RHO, MAT_ARRAYS = 0, 1, 2
BC_DEFAULT, BC_FREE_SURFACE_REF, BC_FREE_SURFACE, BC_WALL = 0, 1, 2, 3
FP64, FP32, FP32_KAHAN = 0, 1, 2

PI = math.pi  # PETSC_PI


def closure_size(order: int) -> int:
    """Closure rows per boundary: 1/4/6 for orders 2/4/6 (op_defs.h:18-30)."""
    return {2: 1, 4: 4, 6: 6}[order]


def min_points(order: int) -> int:
    """Smallest grid the closures fit on: max(2*cls, closure width, 2r+1)."""
    return {2: 3, 4: 8, 6: 12}[order]


@dataclasses.dataclass
class ProblemSpec:
    physics: int
    order: int
    nx: int
    ny: int = 1
    xl: float = -1.0
    xr: float = 1.0
    yl: float = -1.0
    yr: float = 1.0
    ax: float = 0.0
    ay: float = 0.0
    material: int = MAT_UNIT
    bc: int = BC_DEFAULT
    precision: int = FP64
    rho_inv: Optional[np.ndarray] = None  # (ny, nx) float64
    bulk: Optional[np.ndarray] = None     # (ny, nx) float64
    vel_x: Optional[np.ndarray] = None    # (ny, nx) advection a(i) / a(i,j) per point
    vel_y: Optional[np.ndarray] = None    # (ny, nx) advection b(i,j) per point (2D)

    @property
    def components(self) -> int:
        return {ADVECTION_1D: 1, ADVECTION_2D: 1, ACOUSTIC_2D: 3, REFLECTION_1D: 2}[self.physics]

    @property
    def is_1d(self) -> bool:
        return self.physics in (ADVECTION_1D, REFLECTION_1D)

    @property
    def hix(self) -> float:
        return (self.nx - 1) / (self.xr - self.xl)

    @property
    def hiy(self) -> float:
        return 0.0 if self.is_1d else (self.ny - 1) / (self.yr - self.yl)

    @property
    def hx(self) -> float:
        return 1.0 / self.hix

    @property
    def hy(self) -> float:
        return 1.0 if self.is_1d else 1.0 / self.hiy

    @property
    def length(self) -> int:
        return self.nx * self.ny * self.components

    @property
    def dofs(self) -> int:
        return self.length

    def x(self) -> np.ndarray:
        """Grid x_i = xl + i*h as the drivers' initial data use (wave_eq_sim.cpp:225)."""
        return self.xl + np.arange(self.nx) * self.hx

    def y(self) -> np.ndarray:
        return self.yl + np.arange(self.ny) * self.hy

    def with_(self, **kw) -> "ProblemSpec":
        return dataclasses.replace(self, **kw)

    def shape(self):
        return (self.ny, self.nx, self.components)


# --------------------------------------------------------------------------
# Reference demos
# --------------------------------------------------------------------------

def adv1d(n: int, order: int = 4, a: float = 1.5, rstar: float = 0.1, **kw):
    """advection_1D_sim.cpp: x in [-1,1], a(i) = 1.5, Gaussian rstar = 0.1."""
    spec = ProblemSpec(ADVECTION_1D, order, n, 1, -1.0, 1.0, ax=a, **kw)
    return spec, adv1d_exact(spec, 0.0, rstar)


def adv1d_exact(spec: ProblemSpec, t: float, rstar: float = 0.1) -> np.ndarray:
    x = spec.x()
    return np.exp(-(x - spec.ax * t) ** 2 / (rstar * rstar)).astype(np.float64)


def adv2d(n: int, order: int = 4, a: float = 1.5, b: float = -1.0, rstar: float = 0.1, **kw):
    """advection_2D_sim.cpp: [-1,1]^2, a = 1.5, b = -1, Gaussian rstar = 0.1."""
    spec = ProblemSpec(ADVECTION_2D, order, n, n, -1.0, 1.0, -1.0, 1.0, ax=a, ay=b, **kw)
    return spec, adv2d_exact(spec, 0.0, rstar)


def adv2d_exact(spec: ProblemSpec, t: float, rstar: float = 0.1) -> np.ndarray:
    X, Y = np.meshgrid(spec.x(), spec.y())
    g = np.exp(-((X - spec.ax * t) ** 2 + (Y - spec.ay * t) ** 2) / (rstar * rstar))
    return g.reshape(-1).astype(np.float64)


def adv1d_var(n: int, order: int = 4, width: float = 0.2, **kw):
    """1D advection with the spatially varying velocity a(x) = x + 2 (the
    AppCtx a(i) of advection_1D_sim.cpp:19,70 as a field) on [-1, 1]:
    u_t + (x+2) u_x = 0, u0 = exp(-(x/width)^2).  a > 0 everywhere, so the
    left end is the inflow boundary."""
    spec = ProblemSpec(ADVECTION_1D, order, n, 1, -1.0, 1.0, ax=0.0, **kw)
    spec.vel_x = (spec.x() + 2.0).reshape(1, n)
    return spec, adv1d_var_exact(spec, 0.0, width)


def adv1d_var_exact(spec: ProblemSpec, t: float, width: float = 0.2) -> np.ndarray:
    """Characteristics x + 2 = (x0 + 2) e^t: u(x, t) = u0((x + 2) e^-t - 2).
    At the inflow end x = -1 the exact solution is u0(e^-t - 2) <= exp(-25)
    for width 0.2 and t >= 0, so the SAT's zero inflow data is consistent."""
    x0 = (spec.x() + 2.0) * math.exp(-t) - 2.0
    return np.exp(-(x0 / width) ** 2).astype(np.float64)


def adv2d_var(n: int, order: int = 4, width: float = 0.2, **kw):
    """2D advection with a(x, y) = x + 2 and b(x, y) = y - 2 per grid point
    (the AppCtx a(i,j), b(i,j) of advection_2D_sim.cpp:21,74-75 as fields) on
    [-1, 1]^2; inflow through the west and north boundaries."""
    spec = ProblemSpec(ADVECTION_2D, order, n, n, -1.0, 1.0, -1.0, 1.0, ax=0.0, ay=0.0, **kw)
    X, Y = np.meshgrid(spec.x(), spec.y())
    spec.vel_x = X + 2.0
    spec.vel_y = Y - 2.0
    return spec, adv2d_var_exact(spec, 0.0, width)


def adv2d_var_exact(spec: ProblemSpec, t: float, width: float = 0.2) -> np.ndarray:
    """u(x, y, t) = u0((x + 2) e^-t - 2, 2 + (y - 2) e^-t)."""
    X, Y = np.meshgrid(spec.x(), spec.y())
    x0 = (X + 2.0) * math.exp(-t) - 2.0
    y0 = 2.0 + (Y - 2.0) * math.exp(-t)
    return np.exp(-(x0 * x0 + y0 * y0) / (width * width)).reshape(-1).astype(np.float64)


def standing_mode(spec: ProblemSpec, t: float, n: int = 3, m: int = 4) -> np.ndarray:
    """analytic_solution of wave_eq_sim.cpp:211-233 (mode n=3, m=4), AoS [u,v,p]."""
    X, Y = np.meshgrid(spec.x(), spec.y())
    w = math.sqrt(n * n + m * m)
    out = np.empty((spec.ny, spec.nx, 3))
    out[..., 0] = -n * np.cos(n * PI * X) * np.sin(m * PI * Y) * math.sin(PI * w * t) / w
    out[..., 1] = -m * np.sin(n * PI * X) * np.cos(m * PI * Y) * math.sin(PI * w * t) / w
    out[..., 2] = np.sin(PI * n * X) * np.sin(PI * m * Y) * math.cos(PI * w * t)
    return out.reshape(-1)


def wave(n: int, order: int = 4, **kw):
    """demo/wave: rho = 2 + xy with the manufactured forcing, free-surface SAT."""
    spec = ProblemSpec(ACOUSTIC_2D, order, n, n, -1.0, 1.0, -1.0, 1.0,
                       material=MAT_REF_RHO, **kw)
    return spec, standing_mode(spec, 0.0)


def wave_hom(n: int, order: int = 4, **kw):
    """demo/wave_hom: rho = K = 1, free-surface SAT, standing mode initial data."""
    spec = ProblemSpec(ACOUSTIC_2D, order, n, n, -1.0, 1.0, -1.0, 1.0,
                       material=MAT_UNIT, **kw)
    return spec, standing_mode(spec, 0.0)


def reflection(n: int, order: int = 4, **kw):
    """demo/reflection: [u;v]_t = [v_x;u_x] on [-1,1] with u = 0 at both ends."""
    spec = ProblemSpec(REFLECTION_1D, order, n, 1, -1.0, 1.0, **kw)
    return spec, reflection_initial(spec)


def _theta1(x, t, rstar=0.1):
    return np.exp(-(x - t) ** 2 / (rstar * rstar))


def reflection_initial(spec: ProblemSpec) -> np.ndarray:
    """initial_condition of reflection_sim.cpp:212-226."""
    x = spec.x()
    out = np.empty((spec.nx, 2))
    out[:, 0] = -_theta1(x, 0) - _theta1(x, 0)
    out[:, 1] = -_theta1(x, 0) + _theta1(x, 0)
    return out.reshape(-1)


def reflection_exact(spec: ProblemSpec, t: float) -> np.ndarray:
    """analytic_solution of reflection_sim.cpp:191-204 with W = xr - xl, valid for
    t in [1.5, 2.5] (reflection_sim.cpp:71-74)."""
    x = spec.x()
    W = spec.xr - spec.xl
    out = np.empty((spec.nx, 2))
    out[:, 0] = _theta1(x, W - t) + _theta1(x, -W + t)
    out[:, 1] = _theta1(x, W - t) - _theta1(x, -W + t)
    return out.reshape(-1)


def dt_from_cfl(spec: ProblemSpec, cfl: float) -> float:
    """dt = CFL/min(hix, hiy) (wave_eq_sim.cpp:91, advection_1D_sim.cpp:66)."""
    if spec.is_1d:
        return cfl / spec.hix
    return cfl / min(spec.hix, spec.hiy)


# --------------------------------------------------------------------------
# Benchmark workloads
# --------------------------------------------------------------------------

def gaussian_pulse(spec: ProblemSpec, x0: float, y0: float, sigma: float) -> np.ndarray:
    """AoS acoustic state with p = exp(-((x-x0)^2+(y-y0)^2)/sigma^2), u = v = 0
    (the rstar convention of advection_2D_sim.cpp:188-191)."""
    X, Y = np.meshgrid(spec.x(), spec.y())
    out = np.zeros((spec.ny, spec.nx, 3))
    out[..., 2] = np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / (sigma * sigma))
    return out.reshape(-1)


def smooth_field(nx: int, ny: int, rng: np.random.Generator, lo: float = 0.5,
                 hi: float = 2.0, bumps: int = 8) -> np.ndarray:
    """1 + sum of `bumps` seeded Gaussians on the unit square, mapped affinely
    onto [lo, hi]."""
    x = np.linspace(0.0, 1.0, nx)
    y = np.linspace(0.0, 1.0, ny)
    amp = rng.uniform(-1.0, 1.0, bumps)
    cx = rng.uniform(0.0, 1.0, bumps)
    cy = rng.uniform(0.0, 1.0, bumps)
    wid = rng.uniform(0.08, 0.3, bumps)
    f = np.ones((ny, nx))
    for k in range(bumps):
        gx = np.exp(-((x - cx[k]) ** 2) / wid[k] ** 2)
        gy = np.exp(-((y - cy[k]) ** 2) / wid[k] ** 2)
        f += amp[k] * np.outer(gy, gx)
    fmin, fmax = f.min(), f.max()
    if fmax - fmin < 1e-12:
        return np.full((ny, nx), 0.5 * (lo + hi))
    return lo + (hi - lo) * (f - fmin) / (fmax - fmin)


def acoustic_box(n: int, order: int = 4, seed: int = 2025, sigma: float = 0.05,
                 precision: int = FP64, materials: bool = False, bc: int = BC_WALL,
                 material_seed: Optional[int] = None):
    """2D acoustics on the unit square with wall SATs and a seeded Gaussian
    pressure pulse, v = 0.  materials=True draws rho and K from smooth_field
    (two independent seeded streams)."""
    rng = np.random.default_rng(seed)
    x0, y0 = rng.uniform(0.3, 0.7, 2)
    kw = dict(material=MAT_UNIT)
    if materials:
        mrng = np.random.default_rng(seed + 1 if material_seed is None else material_seed)
        rho = smooth_field(n, n, mrng)
        K = smooth_field(n, n, mrng)
        kw = dict(material=MAT_ARRAYS, rho_inv=1.0 / rho, bulk=K)
    spec = ProblemSpec(ACOUSTIC_2D, order, n, n, 0.0, 1.0, 0.0, 1.0, bc=bc,
                       precision=precision, **kw)
    return spec, gaussian_pulse(spec, float(x0), float(y0), sigma)


def baseline_1d_advection():
    """N=201 order-4 1D advection, a=1, Gaussian exp(-(x/0.2)^2), CFL 0.2 to t=2."""
    spec, v0 = adv1d(201, 4, a=1.0, rstar=0.2)
    dt = dt_from_cfl(spec, 0.2)
    return spec, v0, dt, 2.0


def baseline_acoustic(n: int = 1025, order: int = 4, precision: int = FP64, seed: int = 2025):
    """Unit-square acoustics, rho = c = 1, wall SATs, seeded Gaussian (sigma 0.05),
    CFL 0.25."""
    spec, v0 = acoustic_box(n, order, seed=seed, precision=precision)
    return spec, v0, dt_from_cfl(spec, 0.25)


def baseline_reflection(n: int = 4097, order: int = 6, precision: int = FP32_KAHAN, seed: int = 7):
    """Variable rho and K (8 seeded Gaussians each, range 0.5-2), wall SATs,
    Gaussian pulse, CFL 0.25."""
    spec, v0 = acoustic_box(n, order, seed=seed, precision=precision, materials=True)
    return spec, v0, dt_from_cfl(spec, 0.25)


def batched_ensemble(count: int = 256, nmin: int = 129, nmax: int = 1025, order: int = 4,
                     precision: int = FP32, seed: int = 11):
    """Independent acoustic instances, N_k ~ U{nmin..nmax}, each with its own
    seeded materials and Gaussian source."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, count)
    out = []
    for k, n in enumerate(sizes):
        spec, v0 = acoustic_box(int(n), order, seed=seed * 1000 + k, precision=precision,
                                materials=True)
        out.append((spec, v0, dt_from_cfl(spec, 0.25)))
    return out
