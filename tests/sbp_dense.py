"""Dense SBP operators assembled from the coefficient tables (test helper).

The tables are read back from the oracle as doubles and turned into exact
rationals (every coefficient in code/include/sbpops/*.h is a quotient of small
integers).  The dense matrices are built from the SBP structure -- left
closure block, repeated interior stencil, right closure as the negated mirror
(D1_central.h:95-104) -- and are the independent side of the oracle checks.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np


def rat(x: float) -> Fraction:
    f = Fraction(x).limit_denominator(10 ** 7)
    assert float(f) == x, (x, f)
    return f


def tables(ops):
    r, cls, w = ops.r, ops.cls, ops.w
    inner = [rat(ops.inner[k]) for k in range(2 * r + 1)]
    Q = [[rat(ops.Q[a][b]) for b in range(w)] for a in range(cls)]
    hq = [rat(ops.hq[k]) for k in range(cls)]
    hqinv = [rat(ops.hqinv[k]) for k in range(cls)]
    return dict(r=r, cls=cls, w=w, inner=inner, Q=Q, hq=hq, hqinv=hqinv)


def d1_matrix(tab, n, exact=True):
    """n x n first-derivative matrix with unit spacing."""
    r, cls, w = tab["r"], tab["cls"], tab["w"]
    zero = Fraction(0) if exact else 0.0
    conv = (lambda q: q) if exact else float
    D = [[zero] * n for _ in range(n)]
    for i in range(n):
        if i < cls:
            for s in range(w):
                D[i][s] = conv(tab["Q"][i][s])
        elif i >= n - cls:
            m = n - 1 - i
            for s in range(w):
                D[i][n - w + s] = conv(-tab["Q"][m][w - 1 - s])
        else:
            for s in range(2 * r + 1):
                D[i][i - r + s] = conv(tab["inner"][s])
    return D if exact else np.array(D, dtype=np.float64)


def h_diag(tab, n, exact=True):
    cls = tab["cls"]
    one = Fraction(1) if exact else 1.0
    conv = (lambda q: q) if exact else float
    h = [one] * n
    for k in range(cls):
        h[k] = conv(tab["hq"][k])
        h[n - 1 - k] = conv(tab["hq"][k])
    return h if exact else np.array(h)


def hinv_diag(tab, n):
    cls = tab["cls"]
    h = [1.0] * n
    for k in range(cls):
        h[k] = float(tab["hqinv"][k])
        h[n - 1 - k] = float(tab["hqinv"][k])
    return np.array(h)
