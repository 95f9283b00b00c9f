"""Learned coarse-grid corrector: training data from the solver itself, GPU
training with the library's own bf16 MFMA kernels, corrected coarse solves,
and the time-to-accuracy comparison against the plain FP64 path.

Problem family: 2D acoustics on the unit square, wall SATs, order 4, smooth
seeded rho and K in [0.5, 2], seeded Gaussian pressure pulse (sigma 0.03),
v = 0.  Coarse grid 257^2 (dt = h/4); the fine grid 1025^2 nests 4:1, so the
fine solution is compared on the coarse nodes by injection.

A training pair comes from one fine FP64 trajectory: the fine state F_n at a
coarse step n is injected onto the coarse grid, advanced by one coarse FP64
RK4 step (C'), and the target is the injected F_{n+1} minus C'.  The network
sees high-pass features of C' (raw_features) with rho and K, and predicts
that residual.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import time

import numpy as np

from . import capi
from . import problems as P

N_FINE, N_COARSE, RATIO = 1025, 257, 4
SIGMA = 0.03
ORDER = 4
CFL = 0.25
# The correction acts on points at least BAND from every wall: the order-4
# operator's closure rows (4 per wall) keep the plain scheme's values.  Their
# one-step residual holds most of a step's residual energy, but the closure
# error does not build up over the run the way the interior dispersion error
# does, and a loss over every point spends the network on those few rows.
BAND = 4
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "corrector_weights.npz")


def instance(seed: int, n: int, precision: int = P.FP64, sigma: float = None, order: int = ORDER):
    """Seeded instance on an n x n grid; the same seed gives the same fields
    on every nested grid (exact injection) and for every operator order."""
    return P.acoustic_box(n, order, seed=seed, sigma=SIGMA if sigma is None else sigma, materials=True,
                          material_seed=seed + 7, precision=precision)


def _instances(seeds, n, precision, threads=16, order=ORDER):
    with cf.ThreadPoolExecutor(min(threads, len(seeds))) as ex:
        return list(ex.map(lambda s: instance(s, n, precision, order=order), seeds))


def _shift(f, axis, k):
    n = f.shape[axis]
    return np.take(f, np.clip(np.arange(n) + k, 0, n - 1), axis=axis)


def _d4(f, axis):
    """Clamped fourth difference f[i-2] - 4f[i-1] + 6f[i] - 4f[i+1] + f[i+2]."""
    s = lambda k: _shift(f, axis, k)
    return s(-2) - 4 * s(-1) + 6 * f - 4 * s(1) + s(2)


def _dc(f, axis):
    """Clamped central difference (f[i+1] - f[i-1]) / 2."""
    return 0.5 * (_shift(f, axis, 1) - _shift(f, axis, -1))


def raw_features(spec, state_aos):
    """Network inputs per point, float32 [ny][nx][8] (x is axis 1):
      rho^-1 Dx(d4x p), rho^-1 Dy(d4y p), K (Dx(d4x u) + Dy(d4y v)),
      d4 u, d4 v, d4 p (d4 = d4x + d4y), rho, K.
    The first three are, up to a constant, the leading truncation error
    h^4/30 f^(5) of the fourth-order central stencil in each equation, so the
    per-step residual is close to a linear function of the inputs.  The device
    packing kernels (sbp_corr.hip pack_state_tiled_kernel,
    pack_batch_tiled_kernel) evaluate the same expressions."""
    q = np.asarray(state_aos, dtype=np.float64).reshape(spec.ny, spec.nx, 3)
    u, v, p = q[..., 0], q[..., 1], q[..., 2]
    ri = spec.rho_inv if spec.rho_inv is not None else np.ones((spec.ny, spec.nx))
    K = spec.bulk if spec.bulk is not None else np.ones((spec.ny, spec.nx))
    feats = [ri * _dc(_d4(p, 1), 1), ri * _dc(_d4(p, 0), 0), K * (_dc(_d4(u, 1), 1) + _dc(_d4(v, 0), 0)),
             _d4(u, 1) + _d4(u, 0), _d4(v, 1) + _d4(v, 0), _d4(p, 1) + _d4(p, 0), 1.0 / ri, K]
    return np.stack(feats, axis=-1).astype(np.float32)


def solver_pairs(n_instances=2000, seed0=100000, t_train=0.5, chunk=64, snaps=1, log=print):
    """Solver-generated pairs, `snaps` per instance at coarse steps spread
    over [0, t_train): (seed, spec, C', r, n) with C' the coarse FP64 step of
    the injected fine state F_n and r = inj(F_{n+1}) - C', both [ny][nx][3]
    float64.  The snapshots of one instance are M/snaps coarse steps apart
    (M = t_train/dt), starting at a seeded offset."""
    dt_c = P.dt_from_cfl(instance(seed0, N_COARSE)[0], CFL)
    dt_f = dt_c / RATIO
    M = int(round(t_train / dt_c))
    gap = M // snaps
    seeds = np.arange(seed0, seed0 + n_instances)
    first = {int(s): int((i * 7919) % (M if snaps == 1 else gap)) for i, s in enumerate(seeds)}
    snap = {s: [f + m * gap for m in range(snaps)] for s, f in first.items()}
    order = sorted(seeds.tolist(), key=lambda s: first[s])  # similar snapshot times share a batch
    out = []
    t0 = time.time()
    for c0 in range(0, n_instances, chunk):
        group = order[c0:c0 + chunk]
        fine = _instances(group, N_FINE, P.FP64)
        coarse = _instances(group, N_COARSE, P.FP64)
        last = max(snap[s][-1] for s in group)
        F_n, F_n1 = {}, {}  # (k, snapshot) -> injected fine state
        with capi.Batch([f[0] for f in fine]) as fb:
            for k, (_, v0) in enumerate(fine):
                fb.set_state(k, v0)
            for n in range(last + 1):
                ks = [(k, m) for k, s in enumerate(group) for m, t in enumerate(snap[s]) if t == n]
                for km in ks:
                    F_n[km] = fb.get_state_strided(km[0], RATIO)
                fb.step(dt_f, RATIO)
                for km in ks:
                    F_n1[km] = fb.get_state_strided(km[0], RATIO)
        for m in range(snaps):
            with capi.Batch([c[0] for c in coarse]) as cb:
                for k in range(len(group)):
                    cb.set_state(k, F_n.pop((k, m)))
                cb.step(dt_c, 1)
                for k, (spec, _) in enumerate(coarse):
                    Cp = cb.get_state(k).reshape(spec.ny, spec.nx, 3)
                    res = F_n1.pop((k, m)).reshape(spec.ny, spec.nx, 3) - Cp
                    out.append((group[k], spec, Cp, res, snap[group[k]][m]))
        log("training data: %d/%d instances (%.1f s)" % (min(c0 + chunk, n_instances), n_instances, time.time() - t0))
    return out


def wall_distance(ny, nx):
    """Grid distance of every point to the nearest wall, [ny][nx] int."""
    dx = np.minimum(np.arange(nx), nx - 1 - np.arange(nx))
    dy = np.minimum(np.arange(ny), ny - 1 - np.arange(ny))
    return np.minimum(dy[:, None], dx[None, :])


def featurize(pairs, noise=1.0, band=0):
    """(raw inputs, targets) of solver pairs.

    With noise > 0 each pair's coarse state C' is perturbed by seeded white
    noise eta (per component, amplitude U(0, noise) times the pair's own
    residual standard deviation) and the target becomes inj(F_{n+1}) - (C' +
    eta): the network also learns to remove grid-scale content it did not
    make.  Without it the corrected scheme feeds its own high-wavenumber
    output back through the high-pass inputs and grows.  The residual
    standard deviation is taken over the corrected region (points at least
    `band` from every wall)."""
    n = len(pairs)
    H, W = pairs[0][2].shape[:2]
    raws = np.empty((n, H, W, capi.CORR_INPUTS), np.float32)
    tgts = np.empty((n, H, W, 3), np.float32)

    inner = wall_distance(H, W) >= band

    def one(k):
        seed, spec, Cp, res, step = pairs[k]
        if noise > 0:
            rng = np.random.default_rng(seed * 7 + 3 + 1000003 * step)
            amp = noise * rng.uniform() * res[inner].std(axis=0)
            eta = rng.standard_normal(Cp.shape) * amp
            Cp = Cp + eta
            res = res - eta
        raws[k] = raw_features(spec, Cp.reshape(-1))
        tgts[k] = res

    with cf.ThreadPoolExecutor(8) as ex:
        list(ex.map(one, range(n)))
    return raws, tgts


def make_training_data(n_instances=2000, seed0=100000, t_train=0.5, chunk=64, noise=1.0, snaps=1, band=BAND,
                       log=print):
    """solver_pairs followed by featurize."""
    return featurize(solver_pairs(n_instances, seed0, t_train, chunk, snaps, log), noise, band)


def normalization(raw, tgt, band=0):
    """Per-channel input scale (and shift for rho and K, which are not
    high-pass) and per-component output scale, from the points of the
    corrected region (at least `band` from every wall: the clamped
    differences near a wall are far larger than inside)."""
    nin = raw.shape[-1]
    inner = wall_distance(raw.shape[1], raw.shape[2]) >= band
    raw, tgt = raw[:, inner], tgt[:, inner]
    shift = np.zeros(nin, np.float32)
    scale = np.ones(nin, np.float32)
    for c in range(nin):
        x = raw[..., c]
        if c >= 6:
            shift[c] = float(x.mean())
        scale[c] = 1.0 / max(float(x.std()), 1e-12)
    out = np.array([max(float(tgt[..., c].std()), 1e-30) for c in range(3)], np.float32)
    return shift, scale, out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal(32 * capi.CORR_INPUTS * 9) * np.sqrt(2.0 / (9 * capi.CORR_INPUTS)), np.zeros(32),
             rng.standard_normal(32 * 32 * 9) * np.sqrt(2.0 / 288), np.zeros(32),
             rng.standard_normal(3 * 32 * 9) * 1e-3, np.zeros(3)]
    return np.concatenate(parts).astype(np.float32)


def train(raw, tgt, epochs=12, batch=16, lr=2e-3, seed=0, val_frac=0.1, init=None, band=BAND, log=print):
    """Adam on the GPU through sbp_corr_train_step; returns (params,
    normalisation, history).  A random val_frac of the samples is held out.
    init = (params, normalisation) continues from a trained network (its
    normalisation is kept).  The loss covers the points at least `band` from
    every wall."""
    n = raw.shape[0]
    rng = np.random.default_rng(seed)
    order = rng.permutation(n)
    raw, tgt = raw[order], tgt[order]
    nval = max(1, int(n * val_frac))
    ntr = n - nval
    if init is None:
        shift, scale, out = normalization(raw[:ntr], tgt[:ntr], band)
        p0 = init_params(seed)
    else:
        p0, (shift, scale, out) = init
    hist = []
    with capi.Corrector(raw.shape[2], raw.shape[1], batch) as c:
        c.params = p0
        c.set_normalization(shift, scale, out)
        c.band = band
        c.load_dataset(raw, tgt)
        steps_per_epoch = max(1, ntr // batch)
        total = epochs * steps_per_epoch
        it = 0
        t0 = time.time()
        for ep in range(epochs):
            perm = rng.permutation(ntr)
            losses = []
            for s in range(steps_per_epoch):
                lr_t = lr * 0.5 * (1 + np.cos(np.pi * it / total))  # cosine decay
                losses.append(c.train_step(perm[s * batch:(s + 1) * batch], lr_t))
                it += 1
            val = [c.gradients(raw[ntr + k:ntr + k + 1], tgt[ntr + k:ntr + k + 1])[0] for k in range(min(nval, 32))]
            hist.append((ep, float(np.mean(losses)), float(np.mean(val))))
            log("epoch %d train %.4f val %.4f (%.1f s)" % (ep, hist[-1][1], hist[-1][2], time.time() - t0))
        params = c.params
    return params, (shift, scale, out), hist


def save_weights(params, norm, hist, band=BAND, path=WEIGHTS):
    shift, scale, out = norm
    np.savez(path, params=params, shift=shift, scale=scale, out=out, band=np.int32(band), history=np.array(hist))


def load_weights(path=WEIGHTS):
    """(params, (shift, scale, out_scale), band) of a saved network."""
    d = np.load(path, allow_pickle=False)
    band = int(d["band"]) if "band" in d.files else 0
    return d["params"], (d["shift"], d["scale"], d["out"]), band


def _l2(spec, v, ref):
    return float(np.sqrt(spec.hx * spec.hy) * np.linalg.norm(v - ref))


def interp_time_at_error(points, target, rtol=0.01):
    """Time at error `target` on a curve of (time, error) points, piecewise
    linear in (log error, log time) between the two points that bracket it;
    a target within rtol (relative) of the curve's smallest or largest error
    takes that end point's time (the same grid solved another way lands a
    hair off its error); None when nothing brackets it."""
    pts = sorted(points, key=lambda p: p[1])
    for (t_a, e_a), (t_b, e_b) in zip(pts, pts[1:]):
        if e_a <= target <= e_b and e_b > e_a:
            w = (np.log(target) - np.log(e_a)) / (np.log(e_b) - np.log(e_a))
            return float(np.exp(np.log(t_a) + w * (np.log(t_b) - np.log(t_a))))
    for t_e, e_e in (pts[0], pts[-1]):
        if abs(target - e_e) <= rtol * e_e:
            return float(t_e)
    return None


def time_to_accuracy(params, norm, band=BAND, n_eval=32, seed0=900000, t_end=0.5, plain_n=513, bound=1e-4,
                     log=print, equal_error_grids=(321, 385, 449)):
    """Wall time of the ways to solve n_eval held-out instances to t_end, each
    measured by the L2 error against the FP64 1025^2 solution injected at its
    own grid nodes:
      plain         FP64 single-problem path, one instance after another, plain_n^2
      batched_fp32  FP32 ensemble, plain_n^2, one persistent kernel per stage
      learned       FP32 ensemble on 257^2 + the corrector after every step
    the plain path's and the FP32 ensemble's time-versus-error curves over
    257^2 .. plain_n^2 (grids in between measured against their own
    4x-refined FP64 references), the time each needs at L2 = 1e-4 and the
    time the plain path needs at the learned form's mean and worst error;
    and, for comparison, the FP64 ensemble at plain_n^2 (batched_fp64), the
    learned form with FP64 stages (learned_fp64), the plain 257^2 solve
    (the method without the network), and the plain FP64 path and the FP32
    ensemble with the order-6 operators on 257^2."""
    seeds = list(range(seed0, seed0 + n_eval))
    res = {"n_eval": n_eval, "t_end": t_end, "target_l2": 1e-4}
    # fine FP64 reference, batched
    fine = _instances(seeds, N_FINE, P.FP64)
    dt_f = P.dt_from_cfl(fine[0][0], CFL)
    nf = int(round(t_end / dt_f))
    t = time.time()
    with capi.Batch([f[0] for f in fine]) as b:
        for k, (_, v0) in enumerate(fine):
            b.set_state(k, v0)
        b.step(dt_f, nf)
        b.synchronize()
        res["fine_ref_s"] = time.time() - t
        stride_p = (N_FINE - 1) // (plain_n - 1)
        ref_p = [b.get_state_strided(k, stride_p) for k in range(n_eval)]
        ref_c = [b.get_state_strided(k, RATIO) for k in range(n_eval)]
    log("reference: %d instances of %d^2, %d steps (%.2f s)" % (n_eval, N_FINE, nf, res["fine_ref_s"]))

    def run_plain(n, prec, refs, order=ORDER):
        insts = _instances(seeds, n, prec, order=order)
        dt = P.dt_from_cfl(insts[0][0], CFL)
        ns = int(round(t_end / dt))
        errs, tot = [], 0.0
        for (spec, v0), ref in zip(insts, refs):
            with capi.Context(spec) as c:
                c.set_state(v0)
                c.synchronize()
                t = time.time()
                c.step(0.0, dt, ns)
                c.synchronize()
                tot += time.time() - t
                errs.append(_l2(spec, c.get_state(), ref))
        return tot, errs

    tp, ep = run_plain(plain_n, P.FP64, ref_p)
    res["plain_fp64"] = {"grid": plain_n, "wall_s": tp, "max_l2": max(ep), "mean_l2": float(np.mean(ep))}
    log("plain FP64 %d^2: %.3f s, max L2 %.2e" % (plain_n, tp, max(ep)))
    tc, ec = run_plain(N_COARSE, P.FP64, ref_c)
    res["plain_fp64_coarse"] = {"grid": N_COARSE, "wall_s": tc, "max_l2": max(ec), "mean_l2": float(np.mean(ec))}
    log("plain FP64 %d^2 (no network): %.3f s, max L2 %.2e" % (N_COARSE, tc, max(ec)))

    def run_batched(n, prec, refs):
        insts = _instances(seeds, n, prec)
        dt = P.dt_from_cfl(insts[0][0], CFL)
        ns = int(round(t_end / dt))
        with capi.Batch([s for s, _ in insts]) as b:
            for k, (_, v0) in enumerate(insts):
                b.set_state(k, v0)
            b.autotune(dt)
            b.synchronize()
            t = time.time()
            b.step(dt, ns)
            b.synchronize()
            tb = time.time() - t
            eb = [_l2(insts[k][0], b.get_state(k), refs[k]) for k in range(n_eval)]
        return tb, eb

    # batched ensembles at plain_n (FP32: the chosen form; FP64 for comparison)
    for prec, key in ((P.FP32, "batched_fp32"), (P.FP64, "batched_fp64")):
        tb, eb = run_batched(plain_n, prec, ref_p)
        res[key] = {"grid": plain_n, "wall_s": tb, "max_l2": max(eb), "mean_l2": float(np.mean(eb))}
        log("%s %d^2: %.3f s, max L2 %.2e" % (key, plain_n, tb, max(eb)))
    tb = res["batched_fp32"]["wall_s"]
    tbc, ebc = run_batched(N_COARSE, P.FP32, ref_c)

    # learned corrector on the coarse ensemble (FP32 stages: rounding is far
    # below the discretisation error at this grid; FP64 for comparison)
    shift, scale, out = norm
    for prec, key in ((P.FP32, "learned"), (P.FP64, "learned_fp64")):
        insts = _instances(seeds, N_COARSE, prec)
        dt = P.dt_from_cfl(insts[0][0], CFL)
        ns = int(round(t_end / dt))
        with capi.Batch([s for s, _ in insts]) as b, capi.Corrector(N_COARSE, N_COARSE, n_eval) as cor:
            cor.params = params
            cor.set_normalization(shift, scale, out)
            cor.band = band
            for k, (_, v0) in enumerate(insts):
                b.set_state(k, v0)
            b.autotune(dt)  # launch shape of the plain stages (states are restored)
            b.synchronize()
            t = time.time()
            rej = cor.advance_batch(b, dt, ns, bound)
            tl = time.time() - t
            el = [_l2(insts[k][0], b.get_state(k), ref_c[k]) for k in range(n_eval)]
        res[key] = {"grid": N_COARSE, "precision": "fp32" if prec == P.FP32 else "fp64", "wall_s": tl,
                    "max_l2": max(el), "mean_l2": float(np.mean(el)), "rejected_steps": int(np.sum(rej)),
                    "steps": ns * n_eval}
        log("%s %d^2: %.3f s, max L2 %.2e, mean %.2e, rejected %d of %d" % (
            key, N_COARSE, tl, max(el), float(np.mean(el)), int(np.sum(rej)), ns * n_eval))
    tl = res["learned"]["wall_s"]
    res["speedup_batched_fp32"] = tp / tb
    res["speedup_learned"] = tp / tl
    res["gain_learned_vs_plain_coarse_l2"] = float(np.mean(ec) / res["learned"]["mean_l2"])

    # Equal error: the plain FP64 path on grids between 257^2 and plain_n^2,
    # each against its own 4x-refined FP64 reference (they do not nest in the
    # 1025^2 one), gives its time-versus-error curve; the time it needs at the
    # learned form's error (mean and worst instance) is interpolated on it
    # (piecewise linear in log time against log error).
    curve = [(N_COARSE, tc, float(np.mean(ec)), max(ec)), (plain_n, tp, float(np.mean(ep)), max(ep))]
    eb_p = res["batched_fp32"]
    bcurve = [(N_COARSE, tbc, float(np.mean(ebc)), max(ebc)), (plain_n, tb, eb_p["mean_l2"], eb_p["max_l2"])]
    for n in equal_error_grids:
        nfine = RATIO * (n - 1) + 1
        fine_n = _instances(seeds, nfine, P.FP64)
        dtf = P.dt_from_cfl(fine_n[0][0], CFL)
        with capi.Batch([f[0] for f in fine_n]) as b:
            for k, (_, v0) in enumerate(fine_n):
                b.set_state(k, v0)
            b.step(dtf, int(round(t_end / dtf)))
            b.synchronize()
            refs_n = [b.get_state_strided(k, RATIO) for k in range(n_eval)]
        del fine_n
        tn, en = run_plain(n, P.FP64, refs_n)
        curve.append((n, tn, float(np.mean(en)), max(en)))
        log("plain FP64 %d^2 (vs its %d^2 reference): %.3f s, mean L2 %.2e, max %.2e"
            % (n, nfine, tn, float(np.mean(en)), max(en)))
        tbn, ebn = run_batched(n, P.FP32, refs_n)
        bcurve.append((n, tbn, float(np.mean(ebn)), max(ebn)))
        log("batched FP32 %d^2: %.3f s, mean L2 %.2e, max %.2e" % (n, tbn, float(np.mean(ebn)), max(ebn)))
    curve.sort()
    bcurve.sort()
    res["plain_fp64_curve"] = [{"grid": n, "wall_s": t_, "mean_l2": em, "max_l2": ex} for n, t_, em, ex in curve]
    res["batched_fp32_curve"] = [{"grid": n, "wall_s": t_, "mean_l2": em, "max_l2": ex} for n, t_, em, ex in bcurve]
    # time to the target itself: each grid-refined form interpolated at L2 =
    # target on its own curve (mean over the instances, and worst instance);
    # the learned form solves one grid, below the target on the mean
    target_l2 = res["target_l2"]
    for which, col in (("mean", 2), ("max", 3)):
        tp_t = interp_time_at_error([(c[1], c[col]) for c in curve], target_l2)
        tb_t = interp_time_at_error([(c[1], c[col]) for c in bcurve], target_l2)
        res["plain_time_at_target_%s_l2" % which] = tp_t
        res["batched_fp32_time_at_target_%s_l2" % which] = tb_t
        res["speedup_batched_fp32_at_target_%s_l2" % which] = tp_t / tb_t if tp_t and tb_t else None
        reached = res["learned"][which + "_l2"] <= target_l2
        res["speedup_learned_at_target_%s_l2" % which] = tp_t / tl if tp_t and reached else None
    for which, col in (("mean", 2), ("max", 3)):
        target = res["learned"][which + "_l2"]
        t_eq = interp_time_at_error([(c[1], c[col]) for c in curve], target)
        res["plain_time_at_learned_%s_l2" % which] = t_eq
        res["speedup_learned_equal_%s_l2" % which] = t_eq / tl if t_eq else None
        tb_err = res["batched_fp32"][which + "_l2"]
        tb_eq = interp_time_at_error([(c[1], c[col]) for c in curve], tb_err)
        res["speedup_batched_fp32_equal_%s_l2" % which] = tb_eq / tb if tb_eq else None
    log("equal error: plain FP64 needs %.3f s (mean) / %.3f s (max) at the learned form's error: "
        "%.2fx / %.2fx" % (res["plain_time_at_learned_mean_l2"] or -1, res["plain_time_at_learned_max_l2"] or -1,
                           res["speedup_learned_equal_mean_l2"] or -1, res["speedup_learned_equal_max_l2"] or -1))

    # the same comparison with the order-6 operators on the coarse grid: the
    # plain FP64 path one instance after another and the FP32 ensemble
    t6, e6 = run_plain(N_COARSE, P.FP64, ref_c, order=6)
    res["plain_fp64_order6"] = {"grid": N_COARSE, "order": 6, "wall_s": t6, "max_l2": max(e6),
                                "mean_l2": float(np.mean(e6))}
    log("plain FP64 order 6 %d^2: %.3f s, max L2 %.2e, mean %.2e" % (N_COARSE, t6, max(e6), float(np.mean(e6))))
    insts = _instances(seeds, N_COARSE, P.FP32, order=6)
    dt = P.dt_from_cfl(insts[0][0], CFL)
    ns = int(round(t_end / dt))
    with capi.Batch([s for s, _ in insts]) as b:
        for k, (_, v0) in enumerate(insts):
            b.set_state(k, v0)
        b.autotune(dt)
        b.synchronize()
        t = time.time()
        b.step(dt, ns)
        b.synchronize()
        tb6 = time.time() - t
        eb6 = [_l2(insts[k][0], b.get_state(k), ref_c[k]) for k in range(n_eval)]
    res["batched_fp32_order6"] = {"grid": N_COARSE, "order": 6, "wall_s": tb6, "max_l2": max(eb6),
                                  "mean_l2": float(np.mean(eb6))}
    log("batched FP32 order 6 %d^2: %.3f s, max L2 %.2e, mean %.2e" % (N_COARSE, tb6, max(eb6), float(np.mean(eb6))))
    res["speedup_batched_fp32_order6"] = tp / tb6
    res["speedup_batched_fp32_order6_vs_plain_order6"] = t6 / tb6
    return res


def run_learned(nsteps=512, n_eval=16, seed0=900000, bound=1e-4, log=print):
    """Advance the held-out instances nsteps corrected FP32 steps with the
    shipped weights (the learned form of time_to_accuracy, without the
    comparison runs); returns the wall time per step in seconds."""
    params, norm, band = load_weights()
    insts = _instances(list(range(seed0, seed0 + n_eval)), N_COARSE, P.FP32)
    dt = P.dt_from_cfl(insts[0][0], CFL)
    with capi.Batch([s for s, _ in insts]) as b, capi.Corrector(N_COARSE, N_COARSE, n_eval) as cor:
        cor.params = params
        cor.set_normalization(*norm)
        cor.band = band
        for k, (_, v0) in enumerate(insts):
            b.set_state(k, v0)
        b.autotune(dt)
        cfg = b.get_config()
        b.synchronize()
        t = time.time()
        rej = cor.advance_batch(b, dt, nsteps, bound)
        w = time.time() - t
    log("%d corrected steps of %d x %d^2 (FP32): %.1f us per step, %d refused, stage config %s" % (
        nsteps, n_eval, N_COARSE, w / nsteps * 1e6, int(np.sum(rej)), cfg))
    return w / nsteps


def run_batched(nsteps=512, n=513, n_eval=16, seed0=900000, log=print):
    """Advance the held-out instances on n^2 nsteps plain FP32 ensemble steps
    (the batched form of time_to_accuracy, autotuned, without the comparison
    runs); returns the wall time per step in seconds."""
    insts = _instances(list(range(seed0, seed0 + n_eval)), n, P.FP32)
    dt = P.dt_from_cfl(insts[0][0], CFL)
    with capi.Batch([s for s, _ in insts]) as b:
        for k, (_, v0) in enumerate(insts):
            b.set_state(k, v0)
        b.autotune(dt)
        cfg = b.get_config()
        b.synchronize()
        t = time.time()
        b.step(dt, nsteps)
        b.synchronize()
        w = time.time() - t
    log("%d FP32 ensemble steps of %d x %d^2: %.1f us per step, config %s" % (
        nsteps, n_eval, n, w / nsteps * 1e6, cfg))
    return w / nsteps


def main(argv=None):
    """Regenerate the shipped weights (solver pairs, GPU training, time to
    accuracy), or with --run N only advance the held-out instances N
    corrected steps (--batched: N plain FP32 ensemble steps on 513^2).
    python -m hpc-reading-course_amd.corrector [--instances N] [--snaps S]
    [--band B] [--epochs E] [--lr LR] [--noise A] [--out PATH] [--run STEPS [--batched]]"""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--instances", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--noise", type=float, default=1.0)
    ap.add_argument("--snaps", type=int, default=1)
    ap.add_argument("--band", type=int, default=BAND)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--run", type=int, default=0)
    ap.add_argument("--batched", action="store_true")
    a = ap.parse_args(argv)
    say = lambda *x: print(*x, flush=True)
    if a.run:
        (run_batched if a.batched else run_learned)(a.run, log=say)
        return
    raw, tgt = make_training_data(a.instances, noise=a.noise, snaps=a.snaps, band=a.band, log=say)
    params, norm, hist = train(raw, tgt, epochs=a.epochs, lr=a.lr, band=a.band, log=say)
    del raw, tgt
    save_weights(params, norm, hist, band=a.band, path=a.out)
    say(json.dumps(time_to_accuracy(params, norm, band=a.band, n_eval=16, log=say)))


if __name__ == "__main__":
    main()
