"""Adaptive RKF45 on the GPU (sbp_rk45) against the oracle (orc_rk45), and
the fixed-step drivers ts_rk4 (MATCHSTEP) and RK4_custom.

Both run the same algorithm with the same tableau and operation order, and
the FP64 right-hand sides are bit-identical, so the controller takes the same
decisions: the accepted and rejected counts must agree, and the final state
must match to 1e-12 relative in the max norm (the weighted-norm reductions
differ only in summation order).  The controller's rules themselves are
pinned on the CPU (tests/test_ts_rules.py: hand-worked schedules and a dense
recomputation of the first attempts).
"""
import numpy as np
import pytest

from problems_helper import P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    import importlib
    m = importlib.import_module("hpc-reading-course_amd.capi")
    m.load_library()
    return m


CASES = [("wave", 41, 4, 0.2, 1e-6), ("wave_hom", 37, 6, 0.3, 1e-5), ("adv2d", 33, 2, 0.25, 1e-6),
         ("adv1d", 101, 4, 0.5, 1e-7), ("reflection", 81, 6, 0.4, 1e-6)]


@pytest.mark.parametrize("maker,n,order,T,tol", CASES)
def test_rk45_matches_oracle(capi, orc, maker, n, order, T, tol):
    spec, v0 = getattr(P, maker)(n, order)
    print("rk45", maker, n, order, T, tol, flush=True)
    for dt0 in (1e-3, 0.5):  # the second starts with rejections
        ref, acc, rej, dtn = orc.rk45(spec, v0, 0.0, T, dt0, tol, tol)
        with capi.Context(spec) as ctx:
            ctx.set_state(v0)
            a, r, d = ctx.rk45(0.0, T, dt0, tol, tol)
            got = ctx.get_state()
        assert (a, r) == (acc, rej)
        assert d == pytest.approx(dtn, rel=1e-9)
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_ts_rk45_and_errors(capi, orc):
    spec, v0 = P.wave_hom(33, 4)
    v = v0.copy()
    with capi.Context(spec) as ctx:
        a, r = capi.ts_rk45(ctx, 0.1, 0.01, v)
        assert a > 0
        # the resident state keeps stepping with RK4 after an adaptive run
        ctx.step(0.1, 0.001, 20)
        w = ctx.get_state()
    ref, _, _, _ = orc.rk45(spec, v0, 0.0, 0.1, 0.01)
    assert np.max(np.abs(v - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.max(np.abs(w - orc.rk4(spec, ref, 0.1, 0.001, 20))) <= 1e-12 * np.max(np.abs(ref))
    with capi.Context(spec) as ctx:
        with pytest.raises(capi.SbpError) as e:
            ctx.rk45(0.0, 0.1, 0.01)
        assert e.value.code == -5
        ctx.set_state(v0)
        with pytest.raises(capi.SbpError) as e:
            ctx.rk45(0.0, 0.1, 0.01, max_steps=2)
        assert e.value.code == -1
        with pytest.raises(capi.SbpError):
            ctx.rk45(0.0, 0.1, -1.0)
    with capi.Context(spec.with_(precision=P.FP32)) as ctx:
        ctx.set_state(v0)
        with pytest.raises(capi.SbpError) as e:
            ctx.rk45(0.0, 0.1, 0.01)
        assert e.value.code == -1


def test_rk45_consecutive_rejections_match_oracle(capi, orc):
    """Attempts far beyond stability: rejected several times in a row (the
    reject safety applies from the second), then the run continues."""
    spec, _ = P.wave_hom(33, 4)
    v0 = np.random.default_rng(2).standard_normal(spec.length)
    ref, acc, rej, dtn = orc.rk45(spec, v0, 0.0, 0.05, 0.9, 1e-8, 1e-8)
    assert rej >= 2
    with capi.Context(spec) as ctx:
        ctx.set_state(v0)
        a, r, d = ctx.rk45(0.0, 0.05, 0.9, 1e-8, 1e-8)
        got = ctx.get_state()
    assert (a, r) == (acc, rej)
    assert d == pytest.approx(dtn, rel=1e-9)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_rk45_divergence_is_reported(capi, orc):
    """More than 10 rejections of one step and a non-finite state fail with
    SBP_EDIVERGED (-6), as the oracle fails."""
    spec, _ = P.wave_hom(33, 4)
    v0 = np.random.default_rng(3).standard_normal(spec.length)
    _, acc, rej, _, rc = orc.rk45(spec, v0, 0.0, 1.0, 0.1, 1e-150, 0.0, check=False)
    assert rc == -3 and rej == 11
    with capi.Context(spec) as ctx:
        ctx.set_state(v0)
        with pytest.raises(capi.SbpError) as e:
            ctx.rk45(0.0, 1.0, 0.1, 1e-150, 0.0)
        assert e.value.code == -6
        bad = v0.copy()
        bad[7] = np.nan
        ctx.set_state(bad)
        with pytest.raises(capi.SbpError) as e:
            ctx.rk45(0.0, 1.0, 0.1)
        assert e.value.code == -6


@pytest.mark.parametrize("maker,n,order,ratio", [("wave", 41, 4, 10.5), ("wave_hom", 37, 6, 10.005),
                                                 ("adv1d", 101, 4, 37.3), ("reflection", 81, 2, 3.999)])
def test_ts_rk4_matchstep_matches_oracle(capi, orc, maker, n, order, ratio):
    spec, v0 = getattr(P, maker)(n, order)
    dt = P.dt_from_cfl(spec, 0.2)
    T = ratio * dt
    ref, n_ref, tf_ref = orc.ts_rk4(spec, v0, T, dt, threads=8)
    v = v0.copy()
    with capi.Context(spec) as ctx:
        tf = capi.ts_rk4(ctx, T, dt, v)
        n_gpu, _ = ctx.ts_rk4(0.0, 0.0, dt)  # nothing left to do
    assert n_gpu == 0 and tf == tf_ref
    if spec.is_1d:
        assert np.array_equal(v, ref)
    else:
        assert np.max(np.abs(v - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("t_end,dt", [(0.25, 0.1), (0.3, 0.01)])
def test_rk4_custom_matches_oracle(capi, orc, t_end, dt):
    spec, v0 = P.adv1d(101, 4)
    v = v0.copy()
    with capi.Context(spec) as ctx:
        n, h, tf = capi.rk4_custom(ctx, t_end, dt, v)
    tlen = int(np.floor(t_end / dt + 0.5))
    assert n == tlen
    step = t_end / tlen if abs(tlen * dt - t_end) > 1e-14 else dt
    assert h == step
    ref = orc.rk4(spec, v0, 0.0, step, tlen)
    assert np.array_equal(v, ref)
    t = 0.0
    for _ in range(tlen):
        t = t + step
    assert tf == t
