"""FP32 and FP32+Kahan forms against the FP64 oracle.

Bars (stated per test):
* FP32 right-hand side: max-norm error relative to the FP64 oracle <= 1e-5;
* the Kahan state round trip (FP64 -> FP32 value + FP32 compensation -> FP64)
  keeps ~48 bits: relative error <= 1e-13 (plain FP32: <= 6e-8);
* long runs: the Kahan form is at least as accurate as plain FP32 and within
  2e-5 relative max-norm of the FP64 oracle after 2000 steps; over 10^5
  small steps (CFL 0.01) it is at least 10x more accurate than plain FP32
  and within 3e-5.
"""
import numpy as np
import pytest

from problems_helper import P

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def capi():
    import importlib
    m = importlib.import_module("hpc-reading-course_amd.capi")
    m.load_library()
    return m


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("prec", [P.FP32, P.FP32_KAHAN])
@pytest.mark.parametrize("kind", ["acoustic", "materials", "adv2d", "adv1d", "reflection"])
def test_fp32_rhs(capi, orc, order, prec, kind):
    if kind == "acoustic":
        spec, _ = P.wave_hom(83, order, bc=P.BC_WALL)
    elif kind == "materials":
        spec, _ = P.acoustic_box(90, order, materials=True)
    elif kind == "adv2d":
        spec, _ = P.adv2d(70, order)
    elif kind == "adv1d":
        spec, _ = P.adv1d(300, order)
    else:
        spec, _ = P.reflection(300, order)
    spec = spec.with_(precision=prec)
    X = np.linspace(0, 1, spec.length)
    v = np.sin(7 * X) + 0.3 * np.cos(31 * X)
    ref = orc.rhs(spec.with_(precision=P.FP64), 0.0, v)
    with capi.Context(spec) as ctx:
        got = ctx.rhs(0.0, v)
    assert relerr(got, ref) <= 1e-5


@pytest.mark.parametrize("prec,tol", [(P.FP32, 6e-8), (P.FP32_KAHAN, 1e-13)])
def test_state_round_trip(capi, prec, tol):
    spec, v0 = P.acoustic_box(65, 4, precision=prec)
    v = v0 + np.random.default_rng(0).standard_normal(v0.size) * 1e-3
    with capi.Context(spec) as ctx:
        ctx.set_state(v)
        back = ctx.get_state()
    assert relerr(back, v) <= tol


def test_kahan_long_run_accuracy(capi, orc):
    """2000 steps at CFL 0.25 on 129^2 with materials: both FP32 forms stay
    within 2e-5 of the oracle.  At this step size the FP32 rounding of the
    right-hand side dominates and the two forms agree (measured 8.6e-6 and
    8.8e-6, the order between them set by rounding noise), so the bar is
    agreement within 5%; test_kahan_stops_drift_over_1e5_steps shows what
    the compensation buys at small steps."""
    n, order, steps = 129, 4, 2000
    spec64, v0 = P.acoustic_box(n, order, materials=True)
    dt = P.dt_from_cfl(spec64, 0.25)
    ref = orc.rk4(spec64, v0, 0.0, dt, steps, threads=16)
    errs = {}
    for prec in (P.FP32, P.FP32_KAHAN):
        with capi.Context(spec64.with_(precision=prec)) as ctx:
            errs[prec] = relerr(ctx.rk4_advance(v0, 0.0, dt, steps), ref)
    print("fp32 rel err %.3e  kahan rel err %.3e" % (errs[P.FP32], errs[P.FP32_KAHAN]), flush=True)
    assert errs[P.FP32_KAHAN] <= 1.05 * errs[P.FP32]
    assert errs[P.FP32] <= 2e-5 and errs[P.FP32_KAHAN] <= 2e-5


def test_kahan_stops_drift_over_1e5_steps(capi):
    """10^5 RK4 steps at CFL 0.01 of a 65^2 variable-material acoustic
    problem with wall SATs (T = 15.6, dozens of wall-to-wall transits).  The
    reference is the FP64 path (itself within 1e-12 of the oracle).  With
    small steps every increment is small against U, so plain FP32 loses most
    of each increment's low bits in U += increment and the loss builds up
    over the run; the compensated form carries it.  Measured on MI355X:
    FP32 3.7e-5 / 3.9e-4 after 10^4 / 10^5 steps, compensated 2.1e-6 /
    1.2e-5.  (At CFL 0.25 the two agree, about 3e-4 after 10^5 steps: there
    the FP32 rounding of the right-hand side itself dominates, which no
    compensation of the update removes.)"""
    spec64, v0 = P.acoustic_box(65, 4, materials=True, sigma=0.1)
    dt = P.dt_from_cfl(spec64, 0.01)
    marks = (10_000, 100_000)
    errs = {}
    with capi.Context(spec64) as ref:
        ref.set_state(v0)
        refs, done = [], 0
        for m in marks:
            ref.step(done * dt, dt, m - done)
            done = m
            refs.append(ref.get_state())
    for prec in (P.FP32, P.FP32_KAHAN):
        with capi.Context(spec64.with_(precision=prec)) as ctx:
            ctx.set_state(v0)
            done, e = 0, []
            for m, r in zip(marks, refs):
                ctx.step(done * dt, dt, m - done)
                done = m
                e.append(relerr(ctx.get_state(), r))
            errs[prec] = e
    print("rel err at", marks, "fp32", errs[P.FP32], "kahan", errs[P.FP32_KAHAN], flush=True)
    assert errs[P.FP32_KAHAN][1] <= 0.1 * errs[P.FP32][1]
    assert errs[P.FP32_KAHAN][1] <= 3e-5


def test_fp32_energy_non_increasing(capi):
    spec, v0 = P.acoustic_box(257, 6, materials=True, precision=P.FP32_KAHAN)
    dt = P.dt_from_cfl(spec, 0.25)
    with capi.Context(spec) as ctx:
        ctx.set_state(v0)
        E = [ctx.state_energy()]
        for _ in range(8):
            ctx.step(0.0, dt, 25)
            E.append(ctx.state_energy())
    # FP32 rounding allows relative wiggles of order 1e-6 around a flat energy
    assert all(E[k + 1] <= E[k] * (1 + 1e-6) for k in range(len(E) - 1)), E
    assert E[-1] >= 0.95 * E[0]


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("prec", [P.FP32, P.FP32_KAHAN])
def test_fp32_reference_rho_problem(capi, orc, order, prec):
    """demo/wave (rho = 2 + xy with the manufactured forcing, wave_eq_rhs.h:42-64)
    in FP32: the right-hand side at a stage time where the forcing is on is
    within 1e-5 of the FP64 oracle, and 200 RK4 steps within 2e-5."""
    spec, v0 = P.wave(65, order, precision=prec)
    X = np.linspace(0, 1, spec.length)
    v = np.sin(7 * X) + 0.3 * np.cos(31 * X)
    f64 = spec.with_(precision=P.FP64)
    with capi.Context(spec) as ctx:
        got = ctx.rhs(0.03, v)
        dt = P.dt_from_cfl(spec, 0.25)
        run = ctx.rk4_advance(v0, 0.0, dt, 200)
    assert relerr(got, orc.rhs(f64, 0.03, v)) <= 1e-5
    assert relerr(run, orc.rk4(f64, v0, 0.0, dt, 200, threads=8)) <= 2e-5


@pytest.mark.parametrize("prec", [P.FP32, P.FP32_KAHAN])
@pytest.mark.parametrize("schedule", ["one-stage", "two-stage-nontemporal"])
def test_reflection_workload_error_vs_oracle(capi, orc, prec, schedule):
    """The variable-material order-6 workload (rho, K from 8 seeded Gaussians in
    [0.5, 2], wall SATs, Gaussian pulse, CFL 0.25, 2000 RK4 steps) at 257^2, the
    size the FP64 oracle runs in seconds: the FP32 and FP32+Kahan forms within
    2e-5 relative max norm of the oracle (the 4097^2 run in the benchmark states
    its own error against the FP64 GPU path).  The two-stage schedule with
    nontemporal stores is the one the 4097^2 run takes."""
    import os
    spec, v0, dt = P.baseline_reflection(257, precision=prec)
    print("reflection workload 257^2", prec, schedule, flush=True)
    with capi.Context(spec) as ctx:
        if schedule != "one-stage":
            ctx.set_fused(True)
            ctx.set_store_policy(2)
        got = ctx.rk4_advance(v0, 0.0, dt, 2000)
    ref = orc.rk4(spec.with_(precision=P.FP64), v0, 0.0, dt, 2000, threads=min(16, os.cpu_count() or 1))
    err = relerr(got, ref)
    print("relative max error vs FP64 oracle: %.3e" % err, flush=True)
    assert err <= 2e-5
