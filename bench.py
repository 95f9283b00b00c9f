#!/usr/bin/env python3
"""Benchmark of the fused SBP-SAT RK4 path on one MI355X.

Default workload (the throughput target of the path): 2D first-order acoustics
[u,v,p] on the unit square, 4097 x 4097 points, order-4 SBP, wall SATs,
rho = c = 1, seeded Gaussian pressure pulse, CFL 0.25, FP64.  A step is one
classical RK4 step (four fused stage kernels) over the whole grid with the
state resident in HBM.

Prints one JSON line:
  value       DOF-updates/s (nx*ny*3 values advanced one RK4 step per update)
  roofline    HBM roofline of the dominant kernel: the step's longest launch
              (at 4097^2 the pair2d_kernel launch of RK4 stages 3-4), its
              algorithmic bytes per launch (each value it must read or write
              once: launch_bytes) over its mean duration (HIP events on the
              library's stream), against the 8 TB/s peak; `traffic` is the
              same launch's HBM bytes measured by rocprofv3 counters.  The
              whole step's rate on its own bytes, and on the 16-word basis of
              a 3-register RK4 with one launch per stage, are reported beside
              it, and so is `stream_ceiling`: the rate of a flat elementwise
              kernel that reads and writes the launch's planes per point
  cpu_baseline the oracle compiled with the reference's optimisation flags,
              pthreads over a near-square block decomposition, timed on a
              bounded number of steps of the same workload on this host
  secondary   the 1025^2 configuration and the FP32 forms
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the dominant kernel of the default workload (4097^2,
# order 4, rho = K = 1, FP64, two-stage schedule: the pair2d_kernel launch of
# RK4 stages 3-4), from rocprofv3 as FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 KiB
# (the factors that turn the counters of a 1 GiB plain copy in the same runs
# into the bytes it reads and writes), with the commit and the tile height it
# was measured at: profiles/pair2d_acoustic_4097sq_order4_fp64_hbm_counters.md.
MEASURED_TRAFFIC = {(4097, 4, 0, "two-stage", 3): (491316 * 2.0 + 404713 * 1.0) * 1024}
MEASURED_AT = "84c602f"
MEASURED_TILE_ROWS = 40


def pkg(sub):
    import importlib
    return importlib.import_module("hpc-reading-course_amd." + sub)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_dof(precision, materials=False, C=3):
    """Algorithmic HBM bytes per DOF-update of a fused classical RK4 with
    three registers (U, Y, accumulator; one kernel per stage): 9 reads + 7
    writes = 16 words; 19 words (11 + 8) with a Kahan compensation; stored
    rho^-1 and K add 4 stages x 2 words / C.  This is the roofline basis of
    `achieved` and `frac`."""
    w = 8 if precision == 0 else 4
    words = 16 if precision != 2 else 19
    extra = (4 * 2 / C) if materials else 0.0
    return (words + extra) * w


def launch_bytes(spec, fused):
    """Algorithmic HBM bytes of each of a step's four stage slots: every value
    a launch must read or write, once.  Per DOF (point and component), words:
      one stage per launch: U -> Y2 (2), Y2, U -> Y3 (3), Y3, U -> Y4 (3),
                            Y4, U, Y2, Y3 -> U (5; Kahan + compensation R/W 2)
      two-stage (two launches; slots 2 and 4 are empty):
                  launch 1  interior U -> Y3 (2); frame U -> Y2, Y3 (3: Y2 is
                            kept for the frame's stage 4)
                  launch 3  interior Y3, U -> U' (3; Kahan +2; Y2 is
                            recomputed); frame Y3, U, Y2 -> U' (4; Kahan +2)
    Stored rho^-1 and K add 2 words per point to every launch that differentiates
    with them (two-stage: per interior or frame point a launch covers)."""
    w = 8 if spec.precision == 0 else 4
    kahan = spec.precision == 2
    mat = spec.material == 2
    C = spec.components
    order = spec.order
    npts = spec.nx * spec.ny
    if not fused:
        words = [2, 3, 3, 5 + (2 if kahan else 0)]
        return [float(npts * (C * k + (2 if mat else 0)) * w) for k in words]
    cls, r = {2: 1, 4: 4, 6: 6}[order], order // 2
    d = cls + r
    ni = (spec.nx - 2 * d) * (spec.ny - 2 * d)
    nf = npts - ni
    m = 2 if mat else 0
    k2 = 2 if kahan else 0
    b1 = ni * (2 * C + m) + nf * (3 * C + m)
    b3 = ni * ((3 + k2) * C + m) + nf * ((4 + k2) * C + m)
    return [float(b1 * w), 0.0, float(b3 * w), 0.0]


def launch_planes(spec, fused, slot):
    """(planes read, planes written) per interior point of the launch in stage
    slot `slot` (1-4), one plane being one component or material array: the
    byte mix launch_bytes counts, for the flat streaming ceiling."""
    kahan = spec.precision == 2
    m = 2 if spec.material == 2 else 0
    C = spec.components
    if fused:
        if slot == 1:
            return C + m, C
        return (3 if kahan else 2) * C + m, (2 if kahan else 1) * C
    reads = {1: 1, 2: 2, 3: 2, 4: 4 + (1 if kahan else 0)}[slot]
    writes = 1 + (1 if kahan and slot == 4 else 0)
    return reads * C + m, writes * C


TILE_ROWS = 0  # --tile-rows: fixed two-stage tile height instead of sbp_tune_fused (profiling runs)


def gpu_run(spec, v0, dt, steps, warmup, stage_steps=None, fused=None):
    capi = pkg("capi")
    ctx = capi.Context(spec)
    try:
        if fused is not None:
            ctx.set_fused(fused)
        fused = ctx.fused
        ctx.set_state(v0)
        if fused and TILE_ROWS:
            ctx.set_fused(True, TILE_ROWS)
            rows = TILE_ROWS
        else:
            rows = ctx.tune_fused(dt)  # tile height measured on this grid (state kept), before the timed region
        if warmup:
            ctx.step(0.0, dt, warmup)
        ctx.synchronize()
        t0 = time.perf_counter()
        ev_ms = ctx.time_steps(warmup * dt, dt, steps)  # synchronizes at the end
        wall = time.perf_counter() - t0
        stage_ms = ctx.time_stages(dt, stage_steps or max(3, min(steps, 20)))
        energy = ctx.state_energy()
    finally:
        ctx.close()
    return wall, ev_ms, stage_ms, energy, fused, rows


def measure(spec, v0, dt, steps, warmup, fused=None):
    wall, ev_ms, stage_ms, energy, fused, rows = gpu_run(spec, v0, dt, steps, warmup, fused=fused)
    dofs = spec.nx * spec.ny * spec.components
    mat = spec.material == 2
    bpd = bytes_per_dof(spec.precision, mat, spec.components)
    step_kernel_s = float(np.sum(stage_ms)) * 1e-3
    achieved = dofs * bpd / step_kernel_s / 1e9
    lb = launch_bytes(spec, fused)
    k_achieved = sum(lb) / step_kernel_s / 1e9
    dom = int(np.argmax(stage_ms))
    dom_gbs = lb[dom] / (float(stage_ms[dom]) * 1e-3) / 1e9
    return dict(
        wall_s=wall, event_ms=ev_ms, ms_per_step=wall / steps * 1e3,
        dof_updates_per_s=dofs * steps / wall, stage_ms=[float(x) for x in stage_ms],
        bytes_per_dof=bpd, achieved_gbs=achieved, frac=achieved / PEAK_HBM_GBS, energy=energy,
        kernel_bytes_per_dof=sum(lb) / dofs, kernel_achieved_gbs=k_achieved, kernel_frac=k_achieved / PEAK_HBM_GBS,
        launch_bytes=lb, dominant_launch=dom + 1, dominant_gbs=dom_gbs, dominant_frac=dom_gbs / PEAK_HBM_GBS,
        schedule="two-stage" if fused else "one stage per launch", tile_rows=rows)


def batched(count=256, steps=500, warmup=10, prec=1):
    """Ensemble of `count` acoustic instances, N ~ U{129..1025}, order 4, FP32,
    seeded materials and sources; autotuned launch shape; graph-replayed steps."""
    P = pkg("problems")
    capi = pkg("capi")
    t = time.perf_counter()
    ens = P.batched_ensemble(count, precision=prec)
    setup_s = time.perf_counter() - t
    specs = [e[0] for e in ens]
    dts = [e[2] for e in ens]
    with capi.Batch(specs) as b:
        for k, (_, v0, _) in enumerate(ens):
            b.set_state(k, v0)
        choice = b.autotune(dts)
        cfg = b.get_config()
        b.step(dts, warmup)
        b.synchronize()
        t0 = time.perf_counter()
        ev_ms = b.time_steps(dts, steps)
        wall = time.perf_counter() - t0
        e_end = [b.energy(k) for k in range(min(count, 4))]
    dofs = sum(sp.nx * sp.ny * 3 for sp in specs)
    bpd = bytes_per_dof(prec, materials=True)
    achieved = dofs * bpd * steps / (ev_ms * 1e-3) / 1e9
    kbpd = sum(sum(launch_bytes(sp, cfg.get("two_stage", False))) for sp in specs) / dofs
    return {"instances": count, "dofs": dofs, "steps": steps, "precision": ["fp64", "fp32", "fp32_kahan"][prec],
            "dof_updates_per_s": dofs * steps / wall, "ms_per_step": wall / steps * 1e3,
            "event_ms_per_step": ev_ms / steps, "bytes_per_dof": bpd, "achieved_gbs": achieved,
            "frac": achieved / PEAK_HBM_GBS, "kernel_bytes_per_dof": kbpd,
            "kernel_frac": achieved * kbpd / bpd / PEAK_HBM_GBS, "autotune_choice": choice, "config": cfg,
            "setup_s": setup_s, "energy_head": e_end}


def reflection_2d(steps=2000, warmup=0):
    """Variable-material 2D acoustics (rho, K from 8 seeded Gaussians in
    [0.5, 2]), 4097^2, order 6, wall SATs, 2000 RK4 steps: the FP32 and
    FP32+Kahan forms against the FP64 path (which matches the CPU oracle to
    1e-12 in the parity tests), error stated after the full run."""
    P = pkg("problems")
    capi = pkg("capi")
    out, finals = {}, {}
    for prec, name in ((0, "fp64"), (1, "fp32"), (2, "fp32_kahan")):
        spec, v0, dt = P.baseline_reflection(precision=prec)
        with capi.Context(spec) as ctx:
            ctx.set_state(v0)
            ctx.tune_fused(dt)
            ctx.synchronize()
            t0 = time.perf_counter()
            ctx.time_steps(0.0, dt, steps)
            wall = time.perf_counter() - t0
            finals[name] = ctx.get_state()
            stage_ms = ctx.time_stages(dt, 5)
            fused = ctx.fused
        dofs = spec.nx * spec.ny * 3
        bpd = bytes_per_dof(prec, True)
        lb = launch_bytes(spec, fused)
        kbpd = sum(lb) / dofs
        step_s = float(np.sum(stage_ms)) * 1e-3
        # the dominant launch against the flat stream of its own byte mix
        dom = int(np.argmax(stage_ms))
        dom_gbs = lb[dom] / (float(stage_ms[dom]) * 1e-3) / 1e9
        nr, nw = launch_planes(spec, fused, dom + 1)
        stream_gbs = capi.stream_bandwidth(spec.nx * spec.ny * (8 if prec == 0 else 4), nr, nw, 3)
        out[name] = {"dof_updates_per_s": dofs * steps / wall, "ms_per_step": wall / steps * 1e3,
                     "bytes_per_dof": bpd, "frac": dofs * bpd / step_s / 1e9 / PEAK_HBM_GBS,
                     "kernel_bytes_per_dof": kbpd, "kernel_frac": dofs * kbpd / step_s / 1e9 / PEAK_HBM_GBS,
                     "stage_ms": [float(x) for x in stage_ms], "dominant_launch": dom + 1,
                     "dominant_frac": dom_gbs / PEAK_HBM_GBS,
                     "stream_ceiling": {"planes_read": nr, "planes_written": nw, "gbs": stream_gbs,
                                        "frac": dom_gbs / stream_gbs},
                     "schedule": "two-stage" if fused else "one stage per launch"}
        del v0
    ref = finals["fp64"]
    scale = float(np.max(np.abs(ref)))
    h = 1.0 / (4097 - 1)
    for name in ("fp32", "fp32_kahan"):
        d = finals[name] - ref
        out[name]["max_err_rel_vs_fp64"] = float(np.max(np.abs(d)) / scale)
        out[name]["l2_err_vs_fp64"] = float(h * np.linalg.norm(d))
    out["steps"] = steps
    return out


def time_to_l2():
    """Time to L2 error 1e-4 on 16 held-out seeded variable-material acoustic
    instances (T = 0.5, measured against the FP64 1025^2 solve injected on each
    grid): plain FP64 single-problem path vs the FP32 batched ensemble vs the
    learned corrector on 257^2 with the trained weights shipped in the package."""
    C = pkg("corrector")
    if not os.path.exists(C.WEIGHTS):
        return {"skipped": "no trained corrector weights at %s" % C.WEIGHTS}
    params, norm, band = C.load_weights()
    res = C.time_to_accuracy(params, norm, band=band, n_eval=16, log=log)
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(spec, v0, dt, budget_s=12.0):
    from oracle import pyoracle
    threads = min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    v = pyoracle.rk4(spec, v0, 0.0, dt, 2, threads=threads, fast=True)
    per = (time.perf_counter() - t) / 2
    n = int(max(2, min(200, budget_s / max(per, 1e-6))))
    t = time.perf_counter()
    pyoracle.rk4(spec, v, 2 * dt, dt, n, threads=threads, fast=True)
    el = time.perf_counter() - t
    dofs = spec.nx * spec.ny * spec.components
    return {"value": dofs * n / el, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "sample": "%d RK4 steps of the same %dx%d order-%d workload (%.1f s), oracle built with "
                      "-O3 -ffast-math, %d pthreads, near-square block decomposition, on %s (%d logical CPUs)"
                      % (n, spec.nx, spec.ny, spec.order, el, threads, cpu_model(), os.cpu_count() or 0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4097)
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--copy-only", action="store_true",
                    help="only run the 1 GiB plain copy kernel (HBM counter calibration)")
    ap.add_argument("--tta-only", action="store_true",
                    help="only run the time-to-L2-1e-4 comparison (plain FP64, batched FP32, learned)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only run the 256-instance FP32 ensemble (--steps timed steps)")
    ap.add_argument("--reflection-only", action="store_true",
                    help="only run the order-6 variable-material 4097^2 comparison (--steps steps per precision)")
    ap.add_argument("--tile-rows", type=int, default=0,
                    help="two-stage tile height of the main workload (default: tuned on the grid)")
    ap.add_argument("--one-stage", action="store_true",
                    help="run the main workload with one launch per RK4 stage (the stage2d kernels)")
    args = ap.parse_args()
    global TILE_ROWS
    TILE_ROWS = args.tile_rows
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the SBP-SAT path here is built and timed on a single "
              "MI355X (multi-GPU sharding is a design note, not a measured path); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    P = pkg("problems")
    capi = pkg("capi")
    capi.load_library()
    log("device:", capi.device_arch())
    copy_gbs = capi.copy_bandwidth(1 << 30, 5)
    log("plain copy of 1 GiB: %.0f GB/s" % copy_gbs)
    if args.copy_only:
        print(json.dumps({"copy_1GiB_gbs": copy_gbs}))
        return
    if args.tta_only:
        print(json.dumps(time_to_l2()), flush=True)
        return
    if args.batched_only:
        print(json.dumps(batched(steps=args.steps, warmup=args.warmup)), flush=True)
        return
    if args.reflection_only:
        print(json.dumps(reflection_2d(steps=args.steps)), flush=True)
        return

    spec, v0, dt = P.baseline_acoustic(args.n, args.order)
    log("workload %dx%d order %d FP64: %d warmup + %d timed steps" % (spec.nx, spec.ny, spec.order,
                                                                      args.warmup, args.steps))
    main_m = measure(spec, v0, dt, args.steps, args.warmup, fused=False if args.one_stage else None)
    log("main:", json.dumps({k: v for k, v in main_m.items() if k != "stage_ms"}))

    secondary = {}
    if not args.no_secondary:
        s2, v2, dt2 = P.baseline_acoustic(1025, 4)
        m2 = measure(s2, v2, dt2, 1000, 50)
        secondary["acoustic2d_wall_1025sq_order4_fp64"] = {
            "dof_updates_per_s": m2["dof_updates_per_s"], "ms_per_step": m2["ms_per_step"],
            "achieved_gbs": m2["achieved_gbs"], "frac": m2["frac"], "kernel_frac": m2["kernel_frac"],
            "stage_ms": m2["stage_ms"], "schedule": m2["schedule"], "dominant_launch": m2["dominant_launch"],
            "dominant_frac": m2["dominant_frac"]}
        log("1025^2:", secondary["acoustic2d_wall_1025sq_order4_fp64"])
        for prec, name in ((1, "fp32"), (2, "fp32_kahan")):
            s3, v3, dt3 = P.baseline_acoustic(args.n, args.order, precision=prec)
            m3 = measure(s3, v3, dt3, args.steps, args.warmup)
            secondary["acoustic2d_wall_%dsq_order%d_%s" % (args.n, args.order, name)] = {
                "dof_updates_per_s": m3["dof_updates_per_s"], "ms_per_step": m3["ms_per_step"],
                "bytes_per_dof": m3["bytes_per_dof"], "achieved_gbs": m3["achieved_gbs"],
                "frac": m3["frac"], "kernel_bytes_per_dof": m3["kernel_bytes_per_dof"],
                "kernel_frac": m3["kernel_frac"], "stage_ms": m3["stage_ms"], "schedule": m3["schedule"],
                "dominant_launch": m3["dominant_launch"], "dominant_frac": m3["dominant_frac"]}
            log(name, secondary["acoustic2d_wall_%dsq_order%d_%s" % (args.n, args.order, name)])
        s4, v4, dt4 = P.baseline_acoustic(args.n, args.order)
        m4 = measure(s4, v4, dt4, args.steps, args.warmup, fused=False)
        secondary["acoustic2d_wall_%dsq_order%d_fp64_one_stage_per_launch" % (args.n, args.order)] = {
            k: m4[k] for k in ("dof_updates_per_s", "ms_per_step", "achieved_gbs", "frac", "kernel_bytes_per_dof",
                               "kernel_achieved_gbs", "kernel_frac", "stage_ms", "schedule", "dominant_launch",
                               "dominant_frac")}
        log("one stage per launch:", m4)
        del v4
        bm = batched()
        secondary["batched_256x_N129-1025_order4_fp32"] = bm
        log("batched:", bm)
        rf = reflection_2d()
        secondary["reflection2d_4097sq_order6_materials_2000steps"] = rf
        log("reflection 2D:", rf)
        tta = time_to_l2()

    if args.no_secondary:
        tta = None
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(spec, v0, dt)
        log("cpu:", cpu)

    dofs = spec.nx * spec.ny * spec.components
    dom = main_m["dominant_launch"]
    # the dominant launch's byte mix in a flat elementwise kernel (no stencil,
    # halo or tiles): the rate those bytes reach as a plain stream
    nr, nw = launch_planes(spec, main_m["schedule"] == "two-stage", dom)
    plane_bytes = spec.nx * spec.ny * (8 if spec.precision == 0 else 4)
    stream_gbs = capi.stream_bandwidth(plane_bytes, nr, nw, 5)
    log("flat stream, %d planes read + %d written: %.0f GB/s" % (nr, nw, stream_gbs))
    traffic = MEASURED_TRAFFIC.get((spec.nx, spec.order, spec.precision, main_m["schedule"], dom))
    if main_m["schedule"] == "two-stage":
        names = {1: "pair2d_kernel, RK4 stages 1-2 (interior tiles and frame patches)", 2: "(no launch)",
                 3: "pair2d_kernel, RK4 stages 3-4 (interior tiles and frame patches)", 4: "(no launch)"}
    else:
        names = {k: "stage2d_kernel, RK4 stage %d" % k for k in range(1, 5)}
    out = {
        "metric": "DOF-updates/s & % HBM roofline; time to L2 err 1e-4: chosen vs plain FP64",
        "value": main_m["dof_updates_per_s"],
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian pressure pulse, v = 0)",
        "config": {"workload": "acoustic2d_wall_%dsq_order%d_fp64" % (spec.nx, spec.order),
                   "nx": spec.nx, "ny": spec.ny, "components": 3, "order": spec.order,
                   "cfl": 0.25, "bc": "wall SAT", "material": "rho=K=1",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": main_m["dominant_gbs"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": main_m["dominant_frac"], "traffic": traffic,
                     "kernel": names[dom], "schedule": main_m["schedule"],
                     "algorithmic_bytes_per_launch": main_m["launch_bytes"][dom - 1],
                     "mean_launch_ms": main_m["stage_ms"][dom - 1],
                     "bytes_basis": "each value the launch must read or write, once (bench.launch_bytes)",
                     "traffic_unit": "HBM bytes per launch of that kernel (rocprofv3 FETCH_SIZE x 2.0 + "
                                     "WRITE_SIZE x 1.0 KiB, copy-calibrated)",
                     "traffic_measured_at": MEASURED_AT,
                     "traffic_tile_rows": MEASURED_TILE_ROWS, "tile_rows": main_m.get("tile_rows"),
                     "launch_bytes": main_m["launch_bytes"], "stage_ms": main_m["stage_ms"],
                     "step": {"bytes": sum(main_m["launch_bytes"]), "bytes_per_dof": main_m["kernel_bytes_per_dof"],
                              "kernel_ms": float(np.sum(main_m["stage_ms"])),
                              "achieved": main_m["kernel_achieved_gbs"], "frac": main_m["kernel_frac"]},
                     "rate_16word_basis": {"bytes_per_dof": main_m["bytes_per_dof"],
                                           "achieved": main_m["achieved_gbs"], "frac_of_peak": main_m["frac"],
                                           "note": "3-register RK4 with one launch per stage (16 words); "
                                                   "the two-stage schedule needs fewer bytes, so this rate "
                                                   "is not bounded by the peak"},
                     "copy_ceiling_gbs": copy_gbs,
                     "stream_ceiling": {"planes_read": nr, "planes_written": nw, "plane_bytes": plane_bytes,
                                        "gbs": stream_gbs, "frac": main_m["dominant_gbs"] / stream_gbs,
                                        "note": "the dominant launch's byte mix per point in a flat elementwise "
                                                "kernel (sbp_stream_bandwidth); frac is the launch's "
                                                "algorithmic rate over it"}},
        "cpu_baseline": cpu,
        "time_to_l2_1e-4": tta,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
