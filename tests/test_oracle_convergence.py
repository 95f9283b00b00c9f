"""The oracle reproduces the SBP-SAT design convergence rates against the
reference demos' analytic solutions (advection_1D_sim.cpp:171-188,
advection_2D_sim.cpp:188-209, wave_eq_sim.cpp:211-233, reflection_sim.cpp:177-204).

Interior order 2p with boundary closure order p gives a global rate of about
min(2p, p+1)..2p: >= 1.8 / 3.2 / 3.3 on the finest pair for orders 2 / 4 / 6.
"""
import math

import pytest

from problems_helper import P

RATE = {2: 1.8, 4: 3.2, 6: 3.3}
CASES = [
    ("adv1d", lambda s, t: P.adv1d_exact(s, t), [101, 201, 401], 0.4),
    ("reflection", P.reflection_exact, [101, 201, 401], 2.0),
    ("wave_hom", P.standing_mode, [21, 41, 81], 0.2),
    ("wave", P.standing_mode, [21, 41, 81], 0.2),
    ("adv2d", lambda s, t: P.adv2d_exact(s, t), [41, 81, 161], 0.2),
]


def solve_l2(orc, maker, exact, n, order, T, cfl=0.1):
    spec, v0 = getattr(P, maker)(n, order)
    dt = P.dt_from_cfl(spec, cfl)
    v, _, _ = orc.ts_rk4(spec, v0, T, dt)
    return orc.error_norms(spec, v, exact(spec, T))[0]


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_design_convergence(orc, order, case):
    maker, exact, ns, T = case
    errs = [solve_l2(orc, maker, exact, n, order, T) for n in ns]
    rate = math.log2(errs[-2] / errs[-1])
    assert rate >= RATE[order], (maker, order, errs, rate)
    assert errs[-1] < errs[0]
