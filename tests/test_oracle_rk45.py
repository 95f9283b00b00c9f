"""Pins the oracle's adaptive Runge-Kutta-Fehlberg 4(5) (orc_rk45).

* The tableau compiled into the oracle satisfies, in exact rational
  arithmetic, row sums c = A 1, all 17 order conditions up to order 5 for the
  propagated weights, all 8 up to order 4 for the embedded weights, and the
  embedded weights miss order 5 (a slip in any coefficient breaks one).
* One accepted step of a linear autonomous problem equals the method's
  stability polynomial sum_k (b^T A^(k-1) 1) (hL)^k applied to the state,
  with L the dense SBP operator (pins the stage combinations and the use of
  the fifth-order weights).
* Over a run, the distance to a converged fixed-step solution shrinks with
  the tolerance, the run lands on t_end (forced problem), and an oversized
  first step is rejected and recovered from.
"""
from fractions import Fraction as F

import numpy as np
import pytest

from problems_helper import P


def dense_operator(orc, spec):
    """L with F(v) = L v (linear autonomous problem), column by column."""
    n = spec.length
    L = np.empty((n, n))
    e = np.zeros(n)
    for k in range(n):
        e[:] = 0
        e[k] = 1
        L[:, k] = orc.rhs(spec, 0.0, e)
    return L


def tableau(orc):
    c, a, b5, b4 = orc.rkf45_tableau()
    fr = lambda x: F(float(x)).limit_denominator(100000)
    C = [fr(x) for x in c]
    A = [[fr(a[i][j]) if j < 5 else F(0) for j in range(6)] for i in range(6)]
    return C, A, [fr(x) for x in b5], [fr(x) for x in b4], (c, a, b5, b4)


def test_tableau_is_exact_fehlberg(orc):
    C, A, B5, B4, raw = tableau(orc)
    c, a, b5, b4 = raw
    # the doubles are the rationals rounded once
    for x, q in zip(list(c) + list(b5) + list(b4), C + B5 + B4):
        assert float(q) == x
    for i in range(6):
        assert sum(A[i][:i]) == C[i]
    n = 6
    Av = lambda v: [sum(A[i][j] * v[j] for j in range(n)) for i in range(n)]
    dot = lambda u, v: sum(x * y for x, y in zip(u, v))
    one = [F(1)] * n
    c2 = [x * x for x in C]
    c3 = [x ** 3 for x in C]
    Ac = Av(C)
    Ac2 = Av(c2)
    AAc = Av(Ac)
    conds = [  # (vector, expected) for b . vector
        (one, F(1)), (C, F(1, 2)), (c2, F(1, 3)), (Ac, F(1, 6)),
        (c3, F(1, 4)), ([x * y for x, y in zip(C, Ac)], F(1, 8)), (Ac2, F(1, 12)), (AAc, F(1, 24)),
    ]
    order5 = [
        ([x ** 4 for x in C], F(1, 5)), ([x * x * y for x, y in zip(C, Ac)], F(1, 10)),
        ([x * y for x, y in zip(C, Ac2)], F(1, 15)), ([x * y for x, y in zip(C, AAc)], F(1, 30)),
        ([y * y for y in Ac], F(1, 20)), (Av(c3), F(1, 20)), (Av([x * y for x, y in zip(C, Ac)]), F(1, 40)),
        (Av(Ac2), F(1, 60)), (Av(AAc), F(1, 120)),
    ]
    for v, want in conds + order5:
        assert dot(B5, v) == want
    for v, want in conds:
        assert dot(B4, v) == want
    assert any(dot(B4, v) != want for v, want in order5)


def test_one_step_is_the_stability_polynomial(orc):
    spec, v0 = P.wave_hom(11, 4)
    L = dense_operator(orc, spec)
    C, A, B5, _, _ = tableau(orc)
    # gamma_k = b^T A^(k-1) 1
    gam = [F(1)]
    vec = [F(1)] * 6
    for k in range(1, 7):
        gam.append(sum(b * x for b, x in zip(B5, vec)))
        vec = [sum(A[i][j] * vec[j] for j in range(6)) for i in range(6)]
    h = 0.01
    want = np.zeros_like(v0)
    term = v0.copy()
    for k, g in enumerate(gam):
        if k:
            term = h * (L @ term)
        want += float(g) * term
    got, acc, rej, _ = orc.rk45(spec, v0, 0.0, h, h, atol=1e3, rtol=1e3)
    assert (acc, rej) == (1, 0)
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    # and it is not classical RK4: the fifth-order term is present
    assert gam[5] == F(1, 120)


def test_tolerance_controls_the_time_error(orc):
    spec, v0 = P.wave(33, 4)  # forced, time dependent
    T = 0.2
    ref = orc.rk4(spec, v0, 0.0, T / 4000, 4000)  # converged in time
    errs = []
    for tol in (1e-4, 1e-6, 1e-8):
        v, acc, rej, _ = orc.rk45(spec, v0, 0.0, T, 0.01, tol, tol)
        errs.append(np.max(np.abs(v - ref)))
        assert acc > 0
    assert errs[1] < errs[0] / 5 and errs[2] < errs[1] / 5
    assert errs[2] < 1e-6  # global error of a 1e-8 local tolerance over the run


def test_oversized_first_step_is_rejected(orc):
    spec, v0 = P.wave_hom(33, 4)
    a, acc_a, rej_a, _ = orc.rk45(spec, v0, 0.0, 0.2, 1.0, 1e-6, 1e-6)
    b, acc_b, rej_b, _ = orc.rk45(spec, v0, 0.0, 0.2, 1e-3, 1e-6, 1e-6)
    assert rej_a >= 1 and acc_a >= 1
    assert np.max(np.abs(a - b)) < 1e-4
    with pytest.raises(ValueError):
        orc.rk45(spec, v0, 0.0, 0.2, 1e-3, 1e-6, 1e-6, max_steps=3)
