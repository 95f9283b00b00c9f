"""Batched ensemble form: every instance must equal the single-problem path.

Bars: FP32/FP32+Kahan instances are bit-identical to the same problem run
alone through sbp_rk4_advance with one launch per stage (same kernels, same
arithmetic, any tiling), and within 2e-6 relative with two stages per launch
(FMA contraction may differ between the unrolled copies of its row loop);
FP64 instances are bit-identical to the single-problem path in both
schedules and within 1e-12 relative max norm of the CPU oracle; the
autotuner leaves the states untouched.
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


def ensemble(prec, sizes=(20, 37, 129, 300, 64, 257), order=4, materials=True):
    out = []
    for k, n in enumerate(sizes):
        spec, v0 = P.acoustic_box(n, order, seed=100 + k, precision=prec, materials=materials)
        out.append((spec, v0, P.dt_from_cfl(spec, 0.25)))
    return out


@pytest.mark.parametrize("prec", [P.FP32, P.FP32_KAHAN])
@pytest.mark.parametrize("order", [2, 4, 6])
def test_batch_matches_single(capi, prec, order):
    ens = ensemble(prec, order=order)
    steps = 23
    with capi.Batch([e[0] for e in ens]) as b:
        for k, (_, v0, _) in enumerate(ens):
            b.set_state(k, v0)
        b.step([e[2] for e in ens], steps)
        got = [b.get_state(k) for k in range(len(ens))]
    for (spec, v0, dt), g in zip(ens, got):
        with capi.Context(spec) as ctx:
            ref = ctx.rk4_advance(v0, 0.0, dt, steps)
        assert np.array_equal(g, ref), (spec.nx, np.max(np.abs(g - ref)))


def test_batch_fp64_matches_oracle(capi, orc):
    ens = ensemble(P.FP64, sizes=(20, 77, 130), materials=True)
    with capi.Batch([e[0] for e in ens]) as b:
        for k, (_, v0, _) in enumerate(ens):
            b.set_state(k, v0)
        b.step([e[2] for e in ens], 12)
        for k, (spec, v0, dt) in enumerate(ens):
            ref = orc.rk4(spec, v0, 0.0, dt, 12)
            got = b.get_state(k)
            assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
            assert b.energy(k) == pytest.approx(orc.energy(spec, ref), rel=1e-12)


def two_stage_configs(capi):
    out = []
    with capi.Batch([P.acoustic_box(40, 4)[0]]) as b:
        for cfg in range(capi.Batch.num_configs()):
            b.set_config(cfg)
            if b.get_config()["two_stage"]:
                out.append(cfg)
    return out


def test_batch_unit_material_and_configs(capi):
    """Every launch shape gives the same result: bit for bit across the
    one-launch-per-stage shapes; the two-stage shapes (whose row loop is
    unrolled, so a row's arithmetic may be contracted into multiply-adds
    differently depending on its place in the tile) within 2e-6 of each
    other, of the one-stage result and of the single-problem path."""
    ens = ensemble(P.FP32, sizes=(33, 200, 90), materials=False)
    dts = [e[2] for e in ens]
    results = {False: [], True: []}
    for cfg in range(capi.Batch.num_configs()):
        with capi.Batch([e[0] for e in ens]) as b:
            b.set_config(cfg)
            two = b.get_config()["two_stage"]
            for k, (_, v0, _) in enumerate(ens):
                b.set_state(k, v0)
            b.step(dts, 11)
            results[two].append([b.get_state(k) for k in range(len(ens))])
    assert results[True], "no two-stage configuration"
    for r in results[False][1:]:
        for a, c in zip(r, results[False][0]):
            assert np.array_equal(a, c)
    for r in results[True][1:] + [results[False][0]]:
        for a, c in zip(r, results[True][0]):
            assert np.max(np.abs(a - c)) <= 2e-6 * np.max(np.abs(c))
    # the two-stage batch against the two-stage single-problem path
    for (spec, v0, dt), g in zip(ens, results[True][0]):
        with capi.Context(spec) as ctx:
            ctx.set_fused(True)
            ref = ctx.rk4_advance(v0, 0.0, dt, 11)
            assert np.max(np.abs(g - ref)) <= 2e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("prec", [P.FP64, P.FP32, P.FP32_KAHAN])
def test_batch_two_stage_parity_and_switching(capi, orc, prec, order):
    """Two-stage configurations: graph replays (10 steps), odd remainders and
    a switch back to a one-stage configuration mid-run keep every instance
    on the single-problem path; FP64 within 1e-12 of the oracle.  Ragged
    instance sizes, so the instances' frame patches and tiles of each launch's
    work list end mid-wave."""
    print("batch two-stage", prec, order, flush=True)
    ens = ensemble(prec, sizes=(20, 77, 130, 41), order=order)
    dts = [e[2] for e in ens]
    cfgs = two_stage_configs(capi)
    with capi.Batch([e[0] for e in ens]) as b:
        for k, (_, v0, _) in enumerate(ens):
            b.set_state(k, v0)
        b.set_config(cfgs[0])
        b.step(dts, 13)        # 10 from a graph + 3: state in the Y0 pool
        mid = [b.get_state(k) for k in range(len(ens))]
        e_mid = [b.energy(k) for k in range(len(ens))]
        b.step(dts, 21)        # graph of the other parity + 1: back in U
        b.set_config(cfgs[-1])
        b.step(dts, 5)         # odd again
        b.set_config(0)        # one launch per stage: the state moves back first
        b.step(dts, 4)
        got = [b.get_state(k) for k in range(len(ens))]
    for (spec, v0, dt), m, em, g in zip(ens, mid, e_mid, got):
        with capi.Context(spec) as ctx:   # (no state reload: a Kahan state's split would change)
            ctx.set_fused(True)
            ctx.set_state(v0)
            ctx.step(0.0, dt, 13)
            ref_mid = ctx.get_state()
            ctx.step(0.0, dt, 26)
            ctx.set_fused(False)
            ctx.step(0.0, dt, 4)
            ref = ctx.get_state()
            if prec == P.FP64:   # no contraction: bit for bit
                assert np.array_equal(m, ref_mid)
                assert np.array_equal(g, ref)
            else:
                assert np.max(np.abs(m - ref_mid)) <= 2e-6 * np.max(np.abs(ref_mid))
                assert np.max(np.abs(g - ref)) <= 2e-6 * np.max(np.abs(ref))
            assert em == pytest.approx(ctx.energy(ref_mid), rel=1e-5)
        if prec == P.FP64:
            assert np.max(np.abs(g - orc.rk4(spec, v0, 0.0, dt, 43))) <= 1e-12 * np.max(np.abs(g))


def test_batch_two_stage_rejects_small_instances(capi):
    ens = ensemble(P.FP32, sizes=(40, 18), order=6)   # 18 <= 2(CLS+R) = 18
    with capi.Batch([e[0] for e in ens]) as b:
        with pytest.raises(capi.SbpError) as e:
            b.set_config(two_stage_configs(capi)[0])
        assert e.value.code == -1
        assert not b.get_config()["two_stage"]
        for k, (_, v0, _) in enumerate(ens):
            b.set_state(k, v0)
        choice = b.autotune([e[2] for e in ens])   # skips the two-stage shapes
        assert not b.get_config()["two_stage"] and b.get_config()["config"] == choice


def test_batch_autotune_keeps_state(capi):
    ens = ensemble(P.FP32_KAHAN, sizes=(150, 260, 40))
    dts = [e[2] for e in ens]
    with capi.Batch([e[0] for e in ens]) as b:
        for k, (_, v0, _) in enumerate(ens):
            b.set_state(k, v0)
        b.step(dts, 5)
        before = [b.get_state(k) for k in range(len(ens))]
        b.step(dts, 0)
        choice = b.autotune(dts)
        assert 0 <= choice < capi.Batch.num_configs()
        assert b.get_config()["config"] == choice
        b_two_stage = b.get_config()["two_stage"]
        after = [b.get_state(k) for k in range(len(ens))]
        for x, y in zip(before, after):
            assert np.array_equal(x, y)
        b.step(dts, 5)
        final = [b.get_state(k) for k in range(len(ens))]
    for (spec, v0, dt), f in zip(ens, final):
        with capi.Context(spec) as ctx:
            ctx.set_fused(b_two_stage)
            ref = ctx.rk4_advance(v0, 0.0, dt, 10)
            if b_two_stage:
                assert np.max(np.abs(f - ref)) <= 2e-6 * np.max(np.abs(ref))
            else:
                assert np.array_equal(f, ref)


def test_batch_errors(capi):
    with pytest.raises(capi.SbpError):
        capi.Batch([P.wave(20, 4)[0], P.wave(20, 6)[0]])
    ens = ensemble(P.FP32, sizes=(20, 30))
    with capi.Batch([e[0] for e in ens]) as b:
        with pytest.raises(capi.SbpError) as e:
            b.step([1e-3, 1e-3], 1)  # no states loaded yet
        assert e.value.code == -5
        with pytest.raises(capi.SbpError):
            b.set_state(5, ens[0][1])


@pytest.mark.parametrize("policy", [1, 2], ids=["cached", "nontemporal"])
@pytest.mark.parametrize("prec", [P.FP32, P.FP32_KAHAN, P.FP64])
def test_batch_store_policies(capi, orc, prec, policy):
    """Both store builds of the ensemble's two-stage kernels (parts 1 and 3;
    the 256-instance workload runs the nontemporal ones): every instance
    within 2e-6 (FP32; FP64: bit-identical) of the two-stage single-problem
    path; FP64 within 1e-12 of the oracle, FP32 and Kahan within 2e-5 of it
    (the FP32 bar of the single-problem path)."""
    print("batch store policy", prec, policy, flush=True)
    ens = ensemble(prec, sizes=(20, 77, 130, 41, 257))
    dts = [e[2] for e in ens]
    cfgs = two_stage_configs(capi)
    with capi.Batch([e[0] for e in ens]) as b:
        b.set_config(cfgs[0])
        b.set_store_policy(policy)
        for k, (_, v0, _) in enumerate(ens):
            b.set_state(k, v0)
        b.step(dts, 13)
        got = [b.get_state(k) for k in range(len(ens))]
    for (spec, v0, dt), g in zip(ens, got):
        with capi.Context(spec) as ctx:
            ctx.set_fused(True)
            ctx.set_store_policy(policy)
            ref = ctx.rk4_advance(v0, 0.0, dt, 13)
        o = orc.rk4(spec, v0, 0.0, dt, 13)
        if prec == P.FP64:
            assert np.array_equal(g, ref)
            assert np.max(np.abs(g - o)) <= 1e-12 * np.max(np.abs(o))
        else:
            assert np.max(np.abs(g - ref)) <= 2e-6 * np.max(np.abs(ref))
            assert np.max(np.abs(g - o)) <= 2e-5 * np.max(np.abs(o)), (spec.nx, prec)
