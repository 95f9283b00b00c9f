"""Pins the oracle's right-hand sides, boundary terms and RK4.

Independent checks (none restates the oracle's own formulas):
* interior points of every RHS equal a dense Kronecker assembly of the PDE
  from the dense SBP matrices (checks layout, component order, x/y orientation);
* the SBP energy rate v^T H F(v) equals the closed form the SAT terms are
  designed to produce (pins every boundary coefficient and sign):
    advection:            -|a|/2 * (outflow boundary)^2
    acoustics (energy-stable BCs):        0
    reference free surface with rho=2+xy: sum (1-rho_b) u_b p_b terms
    reflection with u=0 at the ends:      0
* one RK4 step equals the 4th-order Taylor polynomial of exp(dt L) with L
  assembled column by column (pins the RK4 combination);
* the forced problem converges at 4th order in time (pins the stage times).
"""
import math

import numpy as np
import pytest

from problems_helper import P
from sbp_dense import d1_matrix, h_diag, tables


def dense(orc, order, n, h):
    tab = tables(orc.ops(order))
    return d1_matrix(tab, n, exact=False) / h, np.array(h_diag(tab, n, exact=False)) * h


def interior_mask(spec):
    m = np.ones((spec.ny, spec.nx), dtype=bool)
    m[:, 0] = m[:, -1] = False
    if not spec.is_1d:
        m[0, :] = m[-1, :] = False
    return m


def rhs_dense_acoustic(orc, spec, v, t=0.0):
    Dx, _ = dense(orc, spec.order, spec.nx, spec.hx)
    Dy, _ = dense(orc, spec.order, spec.ny, spec.hy)
    q = v.reshape(spec.ny, spec.nx, 3)
    u, w, p = q[..., 0], q[..., 1], q[..., 2]
    px = p @ Dx.T
    py = Dy @ p
    ux = u @ Dx.T
    wy = Dy @ w
    if spec.material == P.MAT_ARRAYS:
        ri, K = spec.rho_inv, spec.bulk
    elif spec.material == P.MAT_REF_RHO:
        X = spec.xl + np.arange(spec.nx) / spec.hix
        Y = spec.yl + np.arange(spec.ny) / spec.hiy
        XX, YY = np.meshgrid(X, Y)
        ri, K = 1.0 / (2 + XX * YY), np.ones_like(XX)
    else:
        ri = K = np.ones((spec.ny, spec.nx))
    F = np.empty_like(q)
    F[..., 0] = -ri * px
    F[..., 1] = -ri * py
    F[..., 2] = -K * (ux + wy)
    return F


def energy_weights(spec):
    if spec.material == P.MAT_ARRAYS:
        rho, kinv = 1.0 / spec.rho_inv, 1.0 / spec.bulk
    elif spec.material == P.MAT_REF_RHO:
        X = spec.xl + np.arange(spec.nx) / spec.hix
        Y = spec.yl + np.arange(spec.ny) / spec.hiy
        XX, YY = np.meshgrid(X, Y)
        rho, kinv = 2 + XX * YY, np.ones_like(XX)
    else:
        rho = kinv = np.ones((spec.ny, spec.nx))
    return rho, kinv


def energy_rate(orc, spec, v, F):
    """v^T H W F, the half time-derivative of the discrete energy."""
    _, hx = dense(orc, spec.order, spec.nx, spec.hx)
    if spec.is_1d:
        return float(np.sum(hx[:, None] * v.reshape(spec.nx, -1) * F.reshape(spec.nx, -1)))
    _, hy = dense(orc, spec.order, spec.ny, spec.hy)
    Hw = np.outer(hy, hx)
    if spec.physics == P.ACOUSTIC_2D:
        rho, kinv = energy_weights(spec)
        q, f = v.reshape(spec.ny, spec.nx, 3), F.reshape(spec.ny, spec.nx, 3)
        return float(np.sum(Hw * (rho * (q[..., 0] * f[..., 0] + q[..., 1] * f[..., 1])
                                  + kinv * q[..., 2] * f[..., 2])))
    return float(np.sum(Hw * v.reshape(spec.ny, spec.nx) * F.reshape(spec.ny, spec.nx)))


def material_spec(problems, n, order, bc, seed=3):
    rng = np.random.default_rng(seed)
    rho = problems.smooth_field(n, n, rng)
    K = problems.smooth_field(n, n, rng)
    return P.ProblemSpec(P.ACOUSTIC_2D, order, n, n, 0.0, 1.0, 0.0, 1.0,
                         material=P.MAT_ARRAYS, bc=bc, rho_inv=1.0 / rho, bulk=K)


# ----------------------------------------------------------------------------
@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("kind", ["hom", "ref", "arrays"])
def test_acoustic_interior_matches_dense(orc, problems, order, kind):
    n = max(P.min_points(order) + 3, 14)
    if kind == "hom":
        spec, _ = P.wave_hom(n, order)
    elif kind == "ref":
        spec, _ = P.wave(n, order)
    else:
        spec = material_spec(problems, n, order, P.BC_WALL)
    rng = np.random.default_rng(order)
    v = rng.standard_normal(spec.length)
    t = 0.1  # cos(5 pi t) ~ 6e-17: the manufactured forcing of demo/wave vanishes
    got = orc.rhs(spec, t, v).reshape(spec.ny, spec.nx, 3)
    ref = rhs_dense_acoustic(orc, spec, v)
    m = interior_mask(spec)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got[m] - ref[m])) <= 1e-13 * scale


@pytest.mark.parametrize("order", [2, 4, 6])
def test_advection_interior_matches_dense(orc, order):
    n = P.min_points(order) + 5
    spec, _ = P.adv2d(n, order, a=1.3, b=-0.7)
    Dx, _ = dense(orc, order, n, spec.hx)
    Dy, _ = dense(orc, order, n, spec.hy)
    v = np.random.default_rng(1).standard_normal(spec.length)
    U = v.reshape(n, n)
    ref = -(1.3 * U @ Dx.T + -0.7 * (Dy @ U))
    got = orc.rhs(spec, 0.0, v).reshape(n, n)
    m = interior_mask(spec)
    assert np.max(np.abs(got[m] - ref[m])) <= 1e-13 * np.max(np.abs(ref))
    spec1, _ = P.adv1d(n, order, a=0.9)
    v1 = v[:n]
    ref1 = -0.9 * (Dx @ v1)
    got1 = orc.rhs(spec1, 0.0, v1)
    assert np.max(np.abs(got1[1:-1] - ref1[1:-1])) <= 1e-13 * np.max(np.abs(ref1))


@pytest.mark.parametrize("order", [2, 4, 6])
def test_reflection_interior_matches_dense(orc, order):
    n = P.min_points(order) + 6
    spec, _ = P.reflection(n, order)
    Dx, _ = dense(orc, order, n, spec.hx)
    v = np.random.default_rng(2).standard_normal(spec.length)
    q = v.reshape(n, 2)
    got = orc.rhs(spec, 0.0, v).reshape(n, 2)
    assert np.allclose(got[:, 1], Dx @ q[:, 0], rtol=0, atol=1e-13 * np.abs(Dx @ q[:, 0]).max())
    assert np.allclose(got[1:-1, 0], (Dx @ q[:, 1])[1:-1], rtol=0, atol=1e-13 * np.abs(Dx @ q[:, 1]).max())
    assert got[0, 0] == 0.0 and got[-1, 0] == 0.0


# ----------------------------------------------------------------------------
@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("a", [1.5, -0.8])
def test_advection_1d_energy_rate(orc, order, a):
    spec, _ = P.adv1d(P.min_points(order) + 9, order, a=a)
    v = np.random.default_rng(5).standard_normal(spec.length)
    rate = energy_rate(orc, spec, v, orc.rhs(spec, 0.0, v))
    expect = -a / 2 * v[-1] ** 2 if a > 0 else a / 2 * v[0] ** 2
    assert rate == pytest.approx(expect, rel=1e-11, abs=1e-11)


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("a,b", [(1.5, -1.0), (-0.6, 0.9)])
def test_advection_2d_energy_rate(orc, order, a, b):
    n = P.min_points(order) + 4
    spec, _ = P.adv2d(n, order, a=a, b=b)
    v = np.random.default_rng(6).standard_normal(spec.length)
    rate = energy_rate(orc, spec, v, orc.rhs(spec, 0.0, v))
    U = v.reshape(n, n)
    _, hx = dense(orc, order, n, spec.hx)
    _, hy = dense(orc, order, n, spec.hy)
    xb = U[:, -1] if a > 0 else U[:, 0]
    yb = U[-1, :] if b > 0 else U[0, :]
    expect = -abs(a) / 2 * np.sum(hy * xb ** 2) - abs(b) / 2 * np.sum(hx * yb ** 2)
    assert rate == pytest.approx(expect, rel=1e-11, abs=1e-11)


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("case", ["hom_ref_bc", "arrays_free", "arrays_wall", "hom_wall"])
def test_acoustic_energy_conserved(orc, problems, order, case):
    n = P.min_points(order) + 5
    if case == "hom_ref_bc":
        spec, _ = P.wave_hom(n, order)
    elif case == "hom_wall":
        spec, _ = P.wave_hom(n, order, bc=P.BC_WALL)
    else:
        spec = material_spec(problems, n, order,
                             P.BC_FREE_SURFACE if case == "arrays_free" else P.BC_WALL)
    v = np.random.default_rng(7).standard_normal(spec.length)
    F = orc.rhs(spec, 0.0, v)
    rate = energy_rate(orc, spec, v, F)
    scale = np.sum(np.abs(v)) * np.max(np.abs(F)) * spec.hx * spec.hy
    assert abs(rate) <= 1e-13 * scale


@pytest.mark.parametrize("order", [2, 4, 6])
def test_reference_free_surface_energy_rate(orc, order):
    """demo/wave's free-surface SAT omits rho^-1 (wave_eq_rhs.h:425): the rate is
    sum_j Hy_j[(1-rho_W)u_W p_W - (1-rho_E)u_E p_E] + sum_i Hx_i[(1-rho_S)v_S p_S - (1-rho_N)v_N p_N]."""
    n = P.min_points(order) + 5
    spec, _ = P.wave(n, order)
    v = np.random.default_rng(8).standard_normal(spec.length)
    F = orc.rhs(spec, 0.1, v)
    rate = energy_rate(orc, spec, v, F)
    rho, _ = energy_weights(spec)
    q = v.reshape(n, n, 3)
    _, hx = dense(orc, order, n, spec.hx)
    _, hy = dense(orc, order, n, spec.hy)
    u, w, p = q[..., 0], q[..., 1], q[..., 2]
    expect = (np.sum(hy * ((1 - rho[:, 0]) * u[:, 0] * p[:, 0] - (1 - rho[:, -1]) * u[:, -1] * p[:, -1]))
              + np.sum(hx * ((1 - rho[0, :]) * w[0, :] * p[0, :] - (1 - rho[-1, :]) * w[-1, :] * p[-1, :])))
    # forcing at t=0.1 is ~1e-16 (cos(pi/2) in floating point): included in the tolerance
    assert rate == pytest.approx(expect, rel=1e-10, abs=1e-10)


@pytest.mark.parametrize("order", [2, 4, 6])
def test_reflection_energy_conserved(orc, order):
    spec, _ = P.reflection(P.min_points(order) + 7, order)
    v = np.random.default_rng(9).standard_normal(spec.length)
    v[0] = 0.0
    v[-2] = 0.0  # u = 0 at both ends
    rate = energy_rate(orc, spec, v, orc.rhs(spec, 0.0, v))
    assert abs(rate) <= 1e-12 * np.sum(v * v)


# ----------------------------------------------------------------------------
def assemble(orc, spec, t=0.0):
    n = spec.length
    L = np.empty((n, n))
    e = np.zeros(n)
    for k in range(n):
        e[:] = 0
        e[k] = 1
        L[:, k] = orc.rhs(spec, t, e)
    return L


@pytest.mark.parametrize("maker", ["wave_hom", "adv1d", "adv2d", "reflection"])
def test_rk4_is_taylor_polynomial(orc, maker):
    order = 4
    n = P.min_points(order) + 1
    spec, _ = getattr(P, maker)(n, order)
    L = assemble(orc, spec)
    dt = 0.3 * spec.hx
    v = np.random.default_rng(10).standard_normal(spec.length)
    A = dt * L
    taylor = v.copy()
    term = v.copy()
    for k in range(1, 5):
        term = A @ term / k
        taylor += term
    got = orc.rk4(spec, v, 0.0, dt, 1)
    assert np.max(np.abs(got - taylor)) <= 1e-13 * np.max(np.abs(taylor))
    got2 = orc.rk4(spec, v, 0.0, dt, 2)
    ref2 = np.linalg.matrix_power(np.eye(len(v)) + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24, 2) @ v
    assert np.max(np.abs(got2 - ref2)) <= 1e-12 * np.max(np.abs(ref2))


def test_rk4_stage_times_fourth_order(orc):
    """The forced demo/wave problem is non-autonomous; the stage times
    (0, 1/2, 1/2, 1) give 4th-order temporal convergence."""
    spec, v0 = P.wave(13, 4)
    T = 0.16
    ref = orc.rk4(spec, v0, 0.0, T / 256, 256)
    errs = []
    for m in (8, 16, 32):
        errs.append(np.max(np.abs(orc.rk4(spec, v0, 0.0, T / m, m) - ref)))
    r1 = math.log2(errs[0] / errs[1])
    r2 = math.log2(errs[1] / errs[2])
    assert r1 > 3.6 and r2 > 3.6, (errs, r1, r2)


def test_parallel_rk4_matches_serial(orc):
    spec, v0 = P.wave(41, 4)
    dt = P.dt_from_cfl(spec, 0.2)
    a = orc.rk4(spec, v0, 0.0, dt, 5)
    b = orc.rk4(spec, v0, 0.0, dt, 5, threads=4)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a))
    c = orc.rk4(spec, v0, 0.0, dt, 5, threads=3, fast=True)
    assert np.max(np.abs(a - c)) <= 1e-11 * np.max(np.abs(a))
