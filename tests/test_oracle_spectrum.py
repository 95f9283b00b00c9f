"""Spectrum of the oracle's semi-discrete operators (a stability pin computed
independently of the energy identities): the right-hand side of every
linear, unforced problem is assembled column by column on a tiny grid and
its eigenvalues are taken with numpy.

* Energy-conserving boundary treatments (acoustics with wall, free-surface
  and the reference's free-surface SAT at rho = 1, also with rho/K arrays;
  reflection with u = 0 at both ends) give a purely imaginary spectrum:
  max |Re lambda| <= 1e-10 max |lambda|.
* The upwind advection SATs (constant velocity and the per-point fields,
  1D and 2D) give Re lambda <= 1e-10 max |lambda| with a strictly damped
  part (the outflow boundary removes energy): min Re lambda < -1e-3 max |lambda|.

A boundary term with the wrong sign puts eigenvalues in the right half plane
(checked below: the 1D advection operator with its inflow SAT negated).
"""
import numpy as np
import pytest

from problems_helper import P


def assemble(orc, spec, t=0.0):
    n = spec.length
    L = np.empty((n, n))
    e = np.zeros(n)
    for k in range(n):
        e[:] = 0
        e[k] = 1
        L[:, k] = orc.rhs(spec, t, e)
    return L


def spectrum(orc, spec):
    lam = np.linalg.eigvals(assemble(orc, spec))
    return lam, float(np.max(np.abs(lam)))


def arrays_spec(n, order, bc):
    rng = np.random.default_rng(7)
    rho = P.smooth_field(n, n, rng)
    K = P.smooth_field(n, n, rng)
    return P.ProblemSpec(P.ACOUSTIC_2D, order, n, n, 0.0, 1.0, 0.0, 1.3, material=P.MAT_ARRAYS, bc=bc,
                         rho_inv=1.0 / rho, bulk=K)


CONSERVING = [
    ("wave_hom-wall", lambda o: P.wave_hom(P.min_points(o) + 3, o, bc=P.BC_WALL)[0]),
    ("wave_hom-free", lambda o: P.wave_hom(P.min_points(o) + 3, o, bc=P.BC_FREE_SURFACE)[0]),
    ("wave_hom-ref-free", lambda o: P.wave_hom(P.min_points(o) + 3, o)[0]),
    ("arrays-wall", lambda o: arrays_spec(P.min_points(o) + 2, o, P.BC_WALL)),
    ("reflection", lambda o: P.reflection(3 * P.min_points(o), o)[0]),
]


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("case", CONSERVING, ids=[c[0] for c in CONSERVING])
def test_conserving_spectrum_is_imaginary(orc, order, case):
    spec = case[1](order)
    lam, rad = spectrum(orc, spec)
    assert rad > 0
    assert np.max(np.abs(lam.real)) <= 1e-10 * rad, (case[0], order, np.max(np.abs(lam.real)) / rad)


def adv2d_fields(n, order):
    spec = P.ProblemSpec(P.ADVECTION_2D, order, n, n, -1.0, 1.0, -1.0, 1.0, ax=0.0, ay=0.0)
    X, Y = np.meshgrid(spec.x(), spec.y())
    spec.vel_x, spec.vel_y = X + 2.0, Y - 2.0
    return spec


DISSIPATIVE = [
    ("adv1d", lambda o: P.adv1d(3 * P.min_points(o), o)[0]),
    ("adv1d-negative", lambda o: P.adv1d(3 * P.min_points(o), o, a=-0.7)[0]),
    ("adv1d-field", lambda o: P.adv1d_var(3 * P.min_points(o), o)[0]),
    ("adv2d", lambda o: P.adv2d(P.min_points(o) + 3, o)[0]),
    ("adv2d-fields", lambda o: adv2d_fields(P.min_points(o) + 3, o)),
]


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("case", DISSIPATIVE, ids=[c[0] for c in DISSIPATIVE])
def test_upwind_spectrum_is_damped(orc, order, case):
    spec = case[1](order)
    lam, rad = spectrum(orc, spec)
    assert np.max(lam.real) <= 1e-10 * rad, (case[0], order, np.max(lam.real) / rad)
    assert np.min(lam.real) < -1e-3 * rad


def test_a_wrong_sat_sign_is_unstable(orc):
    """The pin has teeth: the 1D advection operator with its inflow SAT
    negated has an eigenvalue in the right half plane.  The SAT term is the
    operator's L[0, 0] minus the same entry of the operator for -a (whose
    left end is an outflow, with no SAT there), negated back."""
    spec, _ = P.adv1d(30, 4)
    L = assemble(orc, spec)
    spec_out, _ = P.adv1d(30, 4, a=-1.5)
    L_no_sat = -assemble(orc, spec_out)   # -(-a D) = a D without any left SAT
    sat = L[0, 0] - L_no_sat[0, 0]
    assert sat < 0                      # the inflow SAT damps
    bad = L.copy()
    bad[0, 0] -= 2 * sat                # same magnitude, wrong sign
    lam = np.linalg.eigvals(bad)
    assert np.max(lam.real) > 1e-3 * np.max(np.abs(lam))
