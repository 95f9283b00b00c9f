"""Pins the oracle's SBP tables and stencil application.

* The coefficient tables satisfy the SBP property H D + D^T H = diag(-1,0..0,1)
  EXACTLY in rational arithmetic, their inverse-norm table is the exact
  reciprocal of the norm table, constants integrate exactly, and D is exact on
  polynomials of the design degree.  A slip in any single coefficient breaks
  one of these.
* The oracle's C stencil application (orc_d1: left closure, interior, right
  mirror of D1_central.h:57-104) equals the dense matrix product.
"""
from fractions import Fraction

import numpy as np
import pytest

from sbp_dense import d1_matrix, h_diag, tables

ORDERS = [2, 4, 6]


@pytest.mark.parametrize("order", ORDERS)
def test_sbp_identity_exact(orc, order):
    tab = tables(orc.ops(order))
    n = 24
    D = d1_matrix(tab, n)
    H = h_diag(tab, n)
    for a in range(n):
        for b in range(n):
            s = H[a] * D[a][b] + D[b][a] * H[b]
            expect = Fraction(0)
            if a == b == 0:
                expect = Fraction(-1)
            if a == b == n - 1:
                expect = Fraction(1)
            assert s == expect, (order, a, b, s)


@pytest.mark.parametrize("order", ORDERS)
def test_inverse_norm_exact(orc, order):
    tab = tables(orc.ops(order))
    for q, qi in zip(tab["hq"], tab["hqinv"]):
        assert q * qi == 1
    n = 30
    assert sum(h_diag(tab, n)) == n - 1  # constants integrate exactly on [0, n-1]


@pytest.mark.parametrize("order", ORDERS)
def test_polynomial_exactness(orc, order):
    """Interior exact to degree `order`, closures to degree order/2."""
    tab = tables(orc.ops(order))
    n = 28
    D = d1_matrix(tab, n)
    cls = tab["cls"]
    for k in range(order + 1):
        xk = [Fraction(i) ** k for i in range(n)]
        dxk = [k * Fraction(i) ** (k - 1) if k else Fraction(0) for i in range(n)]
        for i in range(n):
            boundary = i < cls or i >= n - cls
            if boundary and k > order // 2:
                continue
            got = sum(D[i][s] * xk[s] for s in range(n))
            assert got == dxk[i], (order, k, i)
    # and the boundary rows are NOT exact one degree higher (the order is sharp)
    k = order // 2 + 1
    xk = [Fraction(i) ** k for i in range(n)]
    errs = [sum(D[i][s] * xk[s] for s in range(n)) - k * Fraction(i) ** (k - 1) for i in range(cls)]
    assert any(e != 0 for e in errs)


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("n", [None, 40])
def test_oracle_d1_matches_dense(orc, order, n):
    tab = tables(orc.ops(order))
    from problems_helper import pmin
    n = n or pmin(order)
    Df = d1_matrix(tab, n, exact=False)
    rng = np.random.default_rng(order * 100 + n)
    v = rng.standard_normal(n)
    hi = 3.7
    ref = hi * (Df @ v)
    got = np.array([orc.d1(order, v, i, hi) for i in range(n)])
    assert np.max(np.abs(got - ref)) <= 1e-13 * hi * np.max(np.abs(v))
