"""Pins the oracle's advection with spatially varying velocity.

The reference's advection demos take the velocity as per-point functions
a(i) and a(i,j), b(i,j) (advection_1D_sim.cpp:19,70; advection_2D_sim.cpp:21,74-75):
the region kernels multiply the derivative by a at each point
(advection_rhs.h:22,36,49,166-393) and every inflow SAT takes tau from the
velocity at its own boundary point (advection_rhs.h:102,116,461-528).

Independent checks:
* the whole right-hand side, boundary rows and corners included, equals a
  dense assembly  F = -diag(a) (D (x) I) u - diag(b) (I (x) D) u + SAT  built
  from the dense SBP matrices, with the SAT weight taken from the dense
  norm H (1/(h q_0)), not from the tabulated inverse weights;
* a constant field given as an array reproduces the scalar-velocity RHS bit
  for bit, and a field that differs from the constant at one boundary point
  changes exactly that point's SAT;
* the solve converges at the design rate to the characteristic solution of
  u_t + (x+2) u_x = 0 (1D) and u_t + (x+2) u_x + (y-2) u_y = 0 (2D).
"""
import math

import numpy as np
import pytest

from problems_helper import P
from sbp_dense import d1_matrix, h_diag, tables


def dense_ops(orc, order, n, h):
    tab = tables(orc.ops(order))
    return d1_matrix(tab, n, exact=False) / h, np.array(h_diag(tab, n, exact=False)) * h


def tau_in(a):   # inflow at the low end: -(a + |a|)/2
    return -0.5 * (a + np.abs(a))


def tau_out(a):  # inflow at the high end: (a - |a|)/2
    return 0.5 * (a - np.abs(a))


def dense_rhs_1d(orc, spec, u):
    D, H = dense_ops(orc, spec.order, spec.nx, spec.hx)
    a = spec.vel_x.reshape(-1)
    F = -a * (D @ u)
    F[0] += 0.5 * tau_in(a[0]) / H[0] * u[0]
    F[-1] += 0.5 * tau_out(a[-1]) / H[-1] * u[-1]
    return F


def dense_rhs_2d(orc, spec, v):
    Dx, Hx = dense_ops(orc, spec.order, spec.nx, spec.hx)
    Dy, Hy = dense_ops(orc, spec.order, spec.ny, spec.hy)
    U = v.reshape(spec.ny, spec.nx)
    A, B = spec.vel_x, spec.vel_y
    F = -(A * (U @ Dx.T) + B * (Dy @ U))
    F[:, 0] += 0.5 * tau_in(A[:, 0]) / Hx[0] * U[:, 0]      # west
    F[0, :] += 0.5 * tau_in(B[0, :]) / Hy[0] * U[0, :]      # south
    F[:, -1] += 0.5 * tau_out(A[:, -1]) / Hx[-1] * U[:, -1]  # east
    F[-1, :] += 0.5 * tau_out(B[-1, :]) / Hy[-1] * U[-1, :]  # north
    return F.reshape(-1)


def random_field(rng, shape, signed):
    f = rng.uniform(0.3, 2.0, shape)
    if signed:
        f *= rng.choice([-1.0, 1.0], shape)
    return f


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("signed", [False, True])
def test_1d_rhs_matches_dense(orc, order, signed):
    n = P.min_points(order) + 9
    rng = np.random.default_rng(10 + order)
    spec, _ = P.adv1d_var(n, order)
    spec.vel_x = random_field(rng, (1, n), signed)
    if signed:  # inflow at both ends
        spec.vel_x[0, 0], spec.vel_x[0, -1] = 1.3, -0.7
    u = rng.standard_normal(n)
    got = orc.rhs(spec, 0.0, u)
    ref = dense_rhs_1d(orc, spec, u)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("signed", [False, True])
def test_2d_rhs_matches_dense(orc, order, signed):
    n = P.min_points(order) + 5
    rng = np.random.default_rng(20 + order)
    spec, _ = P.adv2d_var(n, order)
    spec.vel_x = random_field(rng, (n, n), signed)
    spec.vel_y = random_field(rng, (n, n), signed)
    v = rng.standard_normal(spec.length)
    got = orc.rhs(spec, 0.0, v)
    ref = dense_rhs_2d(orc, spec, v)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("order", [2, 4, 6])
def test_constant_field_equals_scalar_velocity(orc, order):
    n = P.min_points(order) + 4
    v = np.random.default_rng(3).standard_normal(n * n)
    spec, _ = P.adv2d(n, order, a=1.5, b=-1.0)
    ref = orc.rhs(spec, 0.0, v)
    field = spec.with_(vel_x=np.full((n, n), 1.5), vel_y=np.full((n, n), -1.0), ax=0.0, ay=0.0)
    assert np.array_equal(orc.rhs(field, 0.0, v), ref)
    # flip b from -1 to +1 at one north and one south boundary point: only
    # those two points change.  North: the inflow SAT tau_n = (b - |b|)/2 = -1
    # switches off; south: tau_s = -(b + |b|)/2 = -1 switches on.
    field.vel_y = field.vel_y.copy()
    i = n // 2
    field.vel_y[-1, i] = 1.0
    field.vel_y[0, i] = 1.0
    got = orc.rhs(field, 0.0, v).reshape(n, n)
    diff = got - ref.reshape(n, n)
    Dy, Hy = dense_ops(orc, order, n, spec.hy)
    U = v.reshape(n, n)
    dyu = Dy @ U
    expect = np.zeros((n, n))
    expect[-1, i] = -2.0 * dyu[-1, i] - 0.5 * (-1.0) / Hy[-1] * U[-1, i]
    expect[0, i] = -2.0 * dyu[0, i] + 0.5 * (-1.0) / Hy[0] * U[0, i]
    assert np.max(np.abs(diff - expect)) <= 1e-12 * np.max(np.abs(expect))


def solve_l2(orc, maker, exact, n, order, T, cfl=0.1):
    spec, v0 = getattr(P, maker)(n, order)
    vmax = max(float(np.max(np.abs(spec.vel_x))),
               float(np.max(np.abs(spec.vel_y))) if spec.vel_y is not None else 0.0)
    dt = P.dt_from_cfl(spec, cfl) / vmax
    v, _, _ = orc.ts_rk4(spec, v0, T, dt)
    return orc.error_norms(spec, v, exact(spec, T))[0]


RATE = {2: 1.8, 4: 3.2, 6: 3.3}


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("case", [("adv1d_var", P.adv1d_var_exact, [101, 201, 401]),
                                  ("adv2d_var", P.adv2d_var_exact, [41, 81, 161])],
                         ids=["1d", "2d"])
def test_variable_velocity_design_convergence(orc, order, case):
    maker, exact, ns = case
    errs = [solve_l2(orc, maker, exact, n, order, 0.2) for n in ns]
    rate = math.log2(errs[-2] / errs[-1])
    assert rate >= RATE[order], (maker, order, errs, rate)
