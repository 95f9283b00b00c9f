"""HIP path vs the CPU oracle (FP64), through the C-ABI.

Bars (stated per test):
* FP64 right-hand sides (every physics, order, material, boundary kind and
  ragged grid) are BIT-IDENTICAL to the oracle: the FP64 kernels are built
  without FMA contraction and evaluate every stencil, material and SAT term in
  the oracle's (the reference's) operation order.
* 1D RK4 runs are bit-identical too (RK4_custom's combination order).
* 2D RK4 runs accumulate the RK4 weights in a fused register (A = U + dt/6 k1
  + dt/3 k2 + ...) instead of RK4_custom's k1 + 2(k2+k3) + k4: max-norm
  difference relative to the oracle's max norm <= 1e-12.
"""
import os
import sys

import numpy as np
import pytest

from problems_helper import P

pytestmark = pytest.mark.gpu

TOL = 1e-12
THREADS = min(16, os.cpu_count() or 1)


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def capi():
    import importlib
    m = importlib.import_module("hpc-reading-course_amd.capi")
    m.load_library()
    return m


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def material_spec(n, order, bc, seed=3, ny=None):
    ny = ny or n
    rng = np.random.default_rng(seed)
    rho = P.smooth_field(n, ny, rng)
    K = P.smooth_field(n, ny, rng)
    return P.ProblemSpec(P.ACOUSTIC_2D, order, n, ny, 0.0, 1.0, 0.0, 1.3,
                         material=P.MAT_ARRAYS, bc=bc, rho_inv=1.0 / rho, bulk=K)


def cases_2d():
    out = []
    for order in (2, 4, 6):
        m = P.min_points(order)
        for n, ny in ((m, m), (m + 1, m + 3), (37, 29), (300, 45), (531, 270)):
            out.append(("wave_hom", order, n, ny, P.BC_DEFAULT))
            out.append(("wave", order, n, ny, P.BC_DEFAULT))
            out.append(("adv2d", order, n, ny, P.BC_DEFAULT))
            out.append(("arrays", order, n, ny, P.BC_WALL))
        out.append(("arrays", order, 64, 70, P.BC_FREE_SURFACE))
        out.append(("wave_hom", order, 70, 64, P.BC_WALL))
        out.append(("wave", order, 66, 66, P.BC_FREE_SURFACE))
    return out


def make(kind, order, n, ny, bc):
    if kind == "wave_hom":
        spec, _ = P.wave_hom(n, order, bc=bc)
    elif kind == "wave":
        spec, _ = P.wave(n, order, bc=bc)
    elif kind == "adv2d":
        spec, _ = P.adv2d(n, order, a=1.5, b=-1.0)
    else:
        spec = material_spec(n, order, bc, ny=ny)
        return spec
    return spec.with_(ny=ny)


@pytest.mark.parametrize("case", cases_2d(), ids=lambda c: "%s-o%d-%dx%d-bc%d" % c)
def test_rhs_2d(capi, orc, case):
    spec = make(*case)
    say("rhs_2d", case)
    v = np.random.default_rng(hash(case) % 2 ** 32).standard_normal(spec.length)
    t = 0.37
    ref = orc.rhs(spec, t, v)
    with capi.Context(spec) as ctx:
        got = ctx.rhs(t, v)
    assert np.array_equal(got, ref), relerr(got, ref)


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("kind", ["adv1d", "reflection"])
@pytest.mark.parametrize("n", [None, 57, 201, 1000])
def test_rhs_1d(capi, orc, kind, order, n):
    n = n or P.min_points(order)
    spec, _ = getattr(P, kind)(n, order)
    say("rhs_1d", kind, order, n)
    v = np.random.default_rng(n).standard_normal(spec.length)
    ref = orc.rhs(spec, 0.0, v)
    with capi.Context(spec) as ctx:
        got = ctx.rhs(0.0, v)
    assert np.array_equal(got, ref), relerr(got, ref)


RK_CASES = [
    ("wave_hom", 4, 77, 77, P.BC_DEFAULT), ("wave", 4, 61, 61, P.BC_DEFAULT),
    ("wave", 6, 45, 45, P.BC_DEFAULT), ("wave_hom", 2, 50, 41, P.BC_WALL),
    ("adv2d", 6, 90, 90, P.BC_DEFAULT), ("arrays", 4, 300, 280, P.BC_WALL),
    ("arrays", 6, 129, 129, P.BC_FREE_SURFACE), ("arrays", 2, 40, 33, P.BC_WALL),
]


@pytest.mark.parametrize("case", RK_CASES, ids=lambda c: "%s-o%d-%dx%d-bc%d" % c)
@pytest.mark.parametrize("nsteps", [1, 21])
def test_rk4_2d(capi, orc, case, nsteps):
    spec = make(*case)
    say("rk4_2d", case, nsteps)
    rng = np.random.default_rng(5)
    X, Y = np.meshgrid(spec.x(), spec.y())
    v = np.zeros((spec.ny, spec.nx, spec.components))
    for c in range(spec.components):
        v[..., c] = np.exp(-((X - X.mean() - 0.1 * c) ** 2 + (Y - Y.mean()) ** 2) * 20)
    v = v.reshape(-1) + 1e-3 * rng.standard_normal(spec.length)
    dt = P.dt_from_cfl(spec, 0.25)
    ref = orc.rk4(spec, v, 0.1, dt, nsteps, threads=THREADS)
    with capi.Context(spec) as ctx:
        got = ctx.rk4_advance(v, 0.1, dt, nsteps)
    assert relerr(got, ref) <= TOL


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("kind", ["adv1d", "reflection"])
def test_rk4_1d(capi, orc, kind, order):
    spec, v0 = getattr(P, kind)(201, order)
    say("rk4_1d", kind, order)
    dt = P.dt_from_cfl(spec, 0.2)
    ref = orc.rk4(spec, v0, 0.0, dt, 50)
    with capi.Context(spec) as ctx:
        got = ctx.rk4_advance(v0, 0.0, dt, 50)
    assert np.array_equal(got, ref), relerr(got, ref)


def test_config_1d_advection_full(capi, orc):
    """N=201 order 4, a=1, Gaussian exp(-(x/0.2)^2), CFL 0.2 to t=2: 1000 steps."""
    spec, v0, dt, T = P.baseline_1d_advection()
    steps, _ = orc.matchstep_steps(0.0, T, dt)
    assert len(steps) == 1000 and np.allclose(steps, dt, rtol=1e-12, atol=0)
    n = 1000
    say("config 1: 1000 steps")
    ref = orc.rk4(spec, v0, 0.0, dt, n)
    with capi.Context(spec) as ctx:
        got = ctx.rk4_advance(v0, 0.0, dt, n)
        e_gpu = ctx.energy(got)
        v = v0.copy()
        capi.ts_rk4(ctx, T, dt, v)  # the run as ts_rk4 schedules it
    assert np.array_equal(got, ref), relerr(got, ref)
    assert e_gpu == pytest.approx(orc.energy(spec, ref), rel=1e-13)
    assert np.array_equal(v, orc.ts_rk4(spec, v0, T, dt)[0])


def test_config_acoustic_1025_full(capi, orc):
    """1025^2 order 4, rho=c=1, wall SATs, seeded Gaussian, CFL 0.25, 1000 steps."""
    spec, v0, dt = P.baseline_acoustic(1025, 4)
    say("config 2: oracle 1000 steps on", THREADS, "threads")
    ref = orc.rk4(spec, v0, 0.0, dt, 1000, threads=THREADS)
    say("config 2: gpu")
    with capi.Context(spec) as ctx:
        got = ctx.rk4_advance(v0, 0.0, dt, 1000)
        e0 = ctx.energy(v0)
        e1 = ctx.energy(got)
    err = relerr(got, ref)
    say("config 2 rel max err", err)
    assert err <= TOL
    # wall SATs conserve the semi-discrete energy; RK4 dissipates slightly
    assert e1 <= e0 * (1 + 1e-14)
    assert e1 >= 0.99 * e0


def test_energy_matches_oracle(capi, orc):
    for spec in (P.wave(40, 4)[0], material_spec(50, 6, P.BC_WALL), P.adv1d(80, 2)[0],
                 P.reflection(90, 4)[0], P.adv2d(33, 6)[0]):
        v = np.random.default_rng(1).standard_normal(spec.length)
        with capi.Context(spec) as ctx:
            assert ctx.energy(v) == pytest.approx(orc.energy(spec, v), rel=1e-13)
            assert ctx.hnorm(v) == pytest.approx(orc.hnorm(spec, v), rel=1e-13)
            ctx.set_state(v)
            assert ctx.state_energy() == pytest.approx(orc.energy(spec, v), rel=1e-13)


def test_energy_non_increasing_wall(capi):
    spec, v0 = P.acoustic_box(257, 4, materials=True)
    dt = P.dt_from_cfl(spec, 0.25)
    with capi.Context(spec) as ctx:
        ctx.set_state(v0)
        E = [ctx.state_energy()]
        for _ in range(10):
            ctx.step(0.0, dt, 20)
            E.append(ctx.state_energy())
    assert all(E[k + 1] <= E[k] * (1 + 1e-14) for k in range(len(E) - 1)), E
    assert E[-1] > 0.9 * E[0]


def test_tile_rows_do_not_change_results(capi):
    spec, v0, dt = P.baseline_acoustic(300, 6)
    outs = []
    for rows in (1, 7, 16, 64, 300):
        with capi.Context(spec) as ctx:
            ctx.set_tile_rows(rows)
            outs.append(ctx.rk4_advance(v0, 0.0, dt, 3))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_resident_stepping_matches_one_shot(capi):
    spec, v0, dt = P.baseline_acoustic(200, 4)
    with capi.Context(spec) as ctx:
        a = ctx.rk4_advance(v0, 0.0, dt, 40)
        ctx.set_state(v0)
        for _ in range(4):
            ctx.step(0.0, dt, 10)
        b = ctx.get_state()
    assert np.array_equal(a, b)


def test_ts_rk4_matchstep(capi, orc):
    spec, v0 = P.wave(41, 4)
    dt = P.dt_from_cfl(spec, 0.3)
    T = 0.1234
    ref, n, _ = orc.ts_rk4(spec, v0, T, dt, threads=THREADS)
    assert n > 1
    v = v0.copy()
    with capi.Context(spec) as ctx:
        capi.ts_rk4(ctx, T, dt, v)
        dst = np.empty_like(v)
        assert capi.rhs(ctx, 0.2, v, dst) == 0
    assert relerr(v, ref) <= TOL
    assert relerr(dst, orc.rhs(spec, 0.2, v)) <= TOL


def test_errors_are_reported(capi):
    bad = P.ProblemSpec(P.ACOUSTIC_2D, 5, 20, 20)
    with pytest.raises(capi.SbpError) as e:
        capi.Context(bad)
    assert e.value.code == -1
    with pytest.raises(capi.SbpError):
        capi.Context(P.ProblemSpec(P.ACOUSTIC_2D, 4, 5, 20))
    with pytest.raises(capi.SbpError):
        capi.Context(P.ProblemSpec(P.ACOUSTIC_2D, 4, 20, 20, material=P.MAT_ARRAYS))
    spec, v0 = P.wave_hom(20, 4)
    with capi.Context(spec) as ctx:
        with pytest.raises(capi.SbpError) as e:
            ctx.get_state()
        assert e.value.code == -5
    assert capi.device_arch().startswith("gfx950")


def test_bandwidth_probes(capi):
    """The copy and flat streaming probes run, report positive rates below the
    8 TB/s peak, and reject byte mixes outside their instantiated range."""
    cp = capi.copy_bandwidth(64 << 20, 2)
    st = capi.stream_bandwidth(16 << 20, 6, 3, 2)
    assert 0 < cp < 8000 and 0 < st < 8000
    for nr, nw in ((0, 1), (13, 1), (1, 0), (1, 7)):
        with pytest.raises(capi.SbpError):
            capi.stream_bandwidth(1 << 20, nr, nw, 1)


@pytest.mark.parametrize("prec", [P.FP64, P.FP32])
def test_error_norms_on_device(capi, orc, prec):
    """sbp_error_norms (vec_util.h:47-61) reduces on the device in FP64, for
    FP32 contexts too: equal to the oracle's norms to summation order."""
    spec, v0 = P.wave(37, 4)
    spec = spec.with_(material=P.MAT_UNIT, precision=prec)
    rng = np.random.default_rng(3)
    w = v0 + 1e-3 * rng.standard_normal(v0.shape)
    with capi.Context(spec) as ctx:
        l2, mx = ctx.error_norms(v0, w)
    rl2, rmx = orc.error_norms(spec, v0, w)
    assert l2 == pytest.approx(rl2, rel=1e-12)
    assert mx == rmx
