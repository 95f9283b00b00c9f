"""Pins the oracle's discrete energy (orc_energy) and H-norm (orc_hnorm).

Independent sides (none reuses the oracle's weight routine):
* the dense diagonal norm H = h diag(q_0..q_{cls-1}, 1, ..., 1, q_{cls-1}..q_0)
  built from the exact rational tables (sbp_dense.h_diag), contracted with
  numpy: E = sum_ij Hy_j Hx_i [rho (u^2 + v^2) + p^2 / K] for acoustics and
  sum_ij Hy_j Hx_i u^2 otherwise;
* closed forms: H integrates constants exactly (sum H = xr - xl), and the
  order-4 and order-6 norms integrate x^2 and x^2 y^2 exactly (quadrature
  degree 2p - 1 >= 3), so ||1||_H^2, ||x||_H^2 = 2/3 and ||xy||_H^2 = 4/9 on
  [-1, 1]^d are known without any table;
* the reference's get_norm_1D (H_central.h:70-103) run block by block as on
  several ranks (left block: left closure weights, right block: right
  closure weights, centre blocks: h);
* discrimination: a rho <-> rho^-1 swap in the energy weights and a closure
  weight applied unmirrored on the right both move the value far outside
  the tolerances used above, so either slip in the oracle fails this file.

Reference quirk (kept out of the build): on ONE rank get_norm_1D takes only
its "left block" branch (i_start == 0), which weights the right closure
points by h instead of h q_k; the function is never called by the demos.
The build's norm is the symmetric H of the multi-rank path.
"""
import numpy as np
import pytest

from problems_helper import P
from sbp_dense import h_diag, tables


def hvec(orc, order, n, h):
    return np.array(h_diag(tables(orc.ops(order)), n, exact=False)) * h


def material_spec(n, order, seed):
    rng = np.random.default_rng(seed)
    rho = P.smooth_field(n, n + 3, rng)
    K = P.smooth_field(n, n + 3, rng)
    return P.ProblemSpec(P.ACOUSTIC_2D, order, n, n + 3, 0.0, 1.0, 0.0, 1.5, material=P.MAT_ARRAYS,
                         bc=P.BC_WALL, rho_inv=1.0 / rho, bulk=K), rho, K


def dense_energy(orc, spec, v, rho, kinv):
    hx = hvec(orc, spec.order, spec.nx, spec.hx)
    hy = hvec(orc, spec.order, spec.ny, spec.hy)
    q = v.reshape(spec.ny, spec.nx, 3)
    e = rho * (q[..., 0] ** 2 + q[..., 1] ** 2) + kinv * q[..., 2] ** 2
    return float(np.sum(np.outer(hy, hx) * e))


@pytest.mark.parametrize("order", [2, 4, 6])
def test_energy_arrays_matches_dense(orc, order):
    spec, rho, K = material_spec(P.min_points(order) + 11, order, seed=order)
    v = np.random.default_rng(1).standard_normal(spec.length)
    got = orc.energy(spec, v)
    ref = dense_energy(orc, spec, v, rho, 1.0 / K)
    assert got == pytest.approx(ref, rel=1e-13)
    # discrimination: rho <-> rho^-1 swapped, or closure weights not mirrored on the right
    swapped = dense_energy(orc, spec, v, 1.0 / rho, K)
    assert abs(swapped - ref) > 1e-3 * abs(ref)
    tab = tables(orc.ops(order))
    hx = hvec(orc, order, spec.nx, spec.hx)
    hx_bad = hx.copy()
    cls = tab["cls"]
    hx_bad[-cls:] = spec.hx * np.array([float(q) for q in tab["hq"][:cls]])  # unmirrored
    hy = hvec(orc, order, spec.ny, spec.hy)
    q = v.reshape(spec.ny, spec.nx, 3)
    e = rho * (q[..., 0] ** 2 + q[..., 1] ** 2) + q[..., 2] ** 2 / K
    bad = float(np.sum(np.outer(hy, hx_bad) * e))
    if order > 2:  # order 2 has one closure weight: mirroring is the identity
        assert abs(bad - ref) > 1e-6 * abs(ref)


@pytest.mark.parametrize("order", [2, 4, 6])
def test_energy_reference_rho_matches_dense(orc, order):
    spec, v = P.wave(P.min_points(order) + 9, order)
    v = v + np.random.default_rng(2).standard_normal(spec.length)
    X = spec.xl + np.arange(spec.nx) / spec.hix
    Y = spec.yl + np.arange(spec.ny) / spec.hiy
    XX, YY = np.meshgrid(X, Y)
    ref = dense_energy(orc, spec, v, 2 + XX * YY, np.ones_like(XX))
    assert orc.energy(spec, v) == pytest.approx(ref, rel=1e-13)


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("domain", [(0.0, 1.0), (-1.0, 1.0)])
def test_energy_of_constant_state(orc, order, domain):
    """u = v = p = 1 with rho = K = 1: E = 3 |Omega| exactly (H integrates constants)."""
    lo, hi = domain
    n = P.min_points(order) + 20
    spec = P.ProblemSpec(P.ACOUSTIC_2D, order, n, n + 7, lo, hi, lo, hi, bc=P.BC_WALL)
    v = np.ones(spec.length)
    area = (hi - lo) ** 2
    assert orc.energy(spec, v) == pytest.approx(3 * area, rel=1e-14)
    # with a constant material: rho (u^2 + v^2) + p^2 / K
    m = spec.with_(material=P.MAT_ARRAYS, rho_inv=np.full((n + 7, n), 0.5), bulk=np.full((n + 7, n), 4.0))
    assert orc.energy(m, v) == pytest.approx((2 * 2 + 0.25) * area, rel=1e-14)


@pytest.mark.parametrize("order", [4, 6])
def test_hnorm_integrates_polynomials(orc, order):
    n = 41
    spec1, _ = P.adv1d(n, order)
    x = spec1.x()
    assert orc.hnorm(spec1, np.ones(n)) ** 2 == pytest.approx(2.0, rel=1e-14)
    assert orc.hnorm(spec1, x) ** 2 == pytest.approx(2.0 / 3.0, rel=1e-13)
    spec2, _ = P.adv2d(n, order)
    X, Y = np.meshgrid(spec2.x(), spec2.y())
    assert orc.hnorm(spec2, (X * Y).reshape(-1)) ** 2 == pytest.approx(4.0 / 9.0, rel=1e-13)
    # the (two-component) reflection vector sums its components' norms
    s3, _ = P.reflection(n, order)
    q = np.stack([x, np.ones(n)], axis=1).reshape(-1)
    assert orc.hnorm(s3, q) ** 2 == pytest.approx(2.0 / 3.0 + 2.0, rel=1e-13)


def get_norm_1d_blocks(tab, v, h, bounds):
    """get_norm_1D of H_central.h:70-103 summed over rank blocks [i_start, i_end)."""
    N, dofs = v.shape
    cls = tab["cls"]
    q = [float(x) for x in tab["hq"]]
    total = 0.0
    for i_start, i_end in bounds:
        u = 0.0
        for c in range(dofs):
            if i_start == 0:
                for i in range(cls, i_end):
                    u += h * v[i, c] * v[i, c]
                for i in range(cls):
                    u += h * q[i] * v[i, c] * v[i, c]
            elif i_end == N:
                for i in range(i_start, N - cls):
                    u += h * v[i, c] * v[i, c]
                for i in range(N - cls, N):
                    u += h * q[N - i - 1] * v[i, c] * v[i, c]
            else:
                for i in range(i_start, i_end):
                    u += h * v[i, c] * v[i, c]
        total += u
    return np.sqrt(total)


@pytest.mark.parametrize("order", [2, 4, 6])
def test_hnorm_equals_multi_rank_get_norm_1d(orc, order):
    n = 60
    spec, _ = P.reflection(n, order)
    v = np.random.default_rng(order).standard_normal((n, 2))
    tab = tables(orc.ops(order))
    for bounds in ([(0, 30), (30, 60)], [(0, 17), (17, 41), (41, 60)]):
        ref = get_norm_1d_blocks(tab, v, spec.hx, bounds)
        assert orc.hnorm(spec, v.reshape(-1)) == pytest.approx(ref, rel=1e-14)
    # one rank: the reference's left-block branch drops the right closure weights
    serial = get_norm_1d_blocks(tab, v, spec.hx, [(0, n)])
    if order > 2:
        assert abs(serial - orc.hnorm(spec, v.reshape(-1))) > 1e-6 * serial


def test_energy_advection_and_reflection_unweighted(orc):
    spec, _ = P.adv2d(30, 4)
    v = np.random.default_rng(4).standard_normal(spec.length)
    hx = hvec(orc, 4, 30, spec.hx)
    ref = float(np.sum(np.outer(hx, hx) * v.reshape(30, 30) ** 2))
    assert orc.energy(spec, v) == pytest.approx(ref, rel=1e-13)
    assert orc.hnorm(spec, v) ** 2 == pytest.approx(ref, rel=1e-13)
