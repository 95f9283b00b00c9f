#!/usr/bin/env python3
"""Benchmark of the fused SBP-SAT RK4 path on one MI355X.

Default workload (the throughput target of the path): 2D first-order acoustics
[u,v,p] on the unit square, 4097 x 4097 points, order-4 SBP, wall SATs,
rho = c = 1, seeded Gaussian pressure pulse, CFL 0.25, FP64.  A step is one
classical RK4 step (four fused stage kernels) over the whole grid with the
state resident in HBM.

Prints one JSON line:
  value       DOF-updates/s (nx*ny*3 values advanced one RK4 step per update)
  roofline    HBM roofline of the four stage kernels: algorithmic bytes per
              step (13 words per DOF: 9 reads + 4 writes over the four stages)
              over the summed mean stage-kernel durations (HIP events on the
              library's stream)
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
# HBM bytes per DOF-update of the four FP64 stage kernels of the default
# workload (4097^2, order 4, rho = K = 1), from rocprofv3 FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0 (factors from the 1 GiB copy kernel in the same runs):
# profiles/stage2d_acoustic_4097sq_order4_fp64_hbm_counters.md.
MEASURED_TRAFFIC_PER_DOF = {(4097, 4, 0): 107.41}


def pkg(sub):
    import importlib
    return importlib.import_module("hpc-reading-course_amd." + sub)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_dof(precision, materials=False, C=3):
    """Algorithmic HBM bytes per DOF-update of the stage kernels as built:
    9 reads + 4 writes per value and step (U, Y2, Y3, Y4 registers); the Kahan
    form adds the compensation read and write (15 words)."""
    w = 8 if precision == 0 else 4
    words = 13 if precision != 2 else 15
    extra = (4 * 2 / C) if materials else 0.0  # rho^-1 and K read in each of 4 stages
    return (words + extra) * w


def bytes_per_dof_3reg(precision, materials=False, C=3):
    """The 16-word accounting of a fused 3-register RK4 (U, Y, accumulator):
    9 reads + 7 writes; 19 words with Kahan (11 reads + 8 writes)."""
    w = 8 if precision == 0 else 4
    words = 16 if precision != 2 else 19
    extra = (4 * 2 / C) if materials else 0.0
    return (words + extra) * w


def gpu_run(spec, v0, dt, steps, warmup, stage_steps=None):
    capi = pkg("capi")
    ctx = capi.Context(spec)
    try:
        ctx.set_state(v0)
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
    return wall, ev_ms, stage_ms, energy


def measure(spec, v0, dt, steps, warmup):
    wall, ev_ms, stage_ms, energy = gpu_run(spec, v0, dt, steps, warmup)
    dofs = spec.nx * spec.ny * spec.components
    bpd = bytes_per_dof(spec.precision, spec.material == 2, spec.components)
    step_kernel_s = float(np.sum(stage_ms)) * 1e-3
    achieved = dofs * bpd / step_kernel_s / 1e9
    bpd16 = bytes_per_dof_3reg(spec.precision, spec.material == 2, spec.components)
    eq16 = dofs * bpd16 / step_kernel_s / 1e9
    return dict(
        wall_s=wall, event_ms=ev_ms, ms_per_step=wall / steps * 1e3,
        dof_updates_per_s=dofs * steps / wall, stage_ms=[float(x) for x in stage_ms],
        bytes_per_dof=bpd, achieved_gbs=achieved, frac=achieved / PEAK_HBM_GBS, energy=energy,
        bytes_per_dof_3reg=bpd16, achieved_3reg_equiv_gbs=eq16, frac_3reg_equiv=eq16 / PEAK_HBM_GBS)


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
    return {"instances": count, "dofs": dofs, "steps": steps, "precision": ["fp64", "fp32", "fp32_kahan"][prec],
            "dof_updates_per_s": dofs * steps / wall, "ms_per_step": wall / steps * 1e3,
            "event_ms_per_step": ev_ms / steps, "bytes_per_dof": bpd, "achieved_gbs": achieved,
            "frac": achieved / PEAK_HBM_GBS, "autotune_choice": choice, "config": cfg,
            "setup_s": setup_s, "energy_head": e_end}


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
    args = ap.parse_args()
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

    spec, v0, dt = P.baseline_acoustic(args.n, args.order)
    log("workload %dx%d order %d FP64: %d warmup + %d timed steps" % (spec.nx, spec.ny, spec.order,
                                                                      args.warmup, args.steps))
    main_m = measure(spec, v0, dt, args.steps, args.warmup)
    log("main:", json.dumps({k: v for k, v in main_m.items() if k != "stage_ms"}))

    secondary = {}
    if not args.no_secondary:
        s2, v2, dt2 = P.baseline_acoustic(1025, 4)
        m2 = measure(s2, v2, dt2, 1000, 50)
        secondary["acoustic2d_wall_1025sq_order4_fp64"] = {
            "dof_updates_per_s": m2["dof_updates_per_s"], "ms_per_step": m2["ms_per_step"],
            "achieved_gbs": m2["achieved_gbs"], "stage_ms": m2["stage_ms"]}
        log("1025^2:", secondary["acoustic2d_wall_1025sq_order4_fp64"])
        for prec, name in ((1, "fp32"), (2, "fp32_kahan")):
            s3, v3, dt3 = P.baseline_acoustic(args.n, args.order, precision=prec)
            m3 = measure(s3, v3, dt3, args.steps, args.warmup)
            secondary["acoustic2d_wall_%dsq_order%d_%s" % (args.n, args.order, name)] = {
                "dof_updates_per_s": m3["dof_updates_per_s"], "ms_per_step": m3["ms_per_step"],
                "bytes_per_dof": m3["bytes_per_dof"], "achieved_gbs": m3["achieved_gbs"],
                "frac": m3["frac"], "stage_ms": m3["stage_ms"]}
            log(name, secondary["acoustic2d_wall_%dsq_order%d_%s" % (args.n, args.order, name)])
        bm = batched()
        secondary["batched_256x_N129-1025_order4_fp32"] = bm
        log("batched:", bm)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(spec, v0, dt)
        log("cpu:", cpu)

    traffic_dof = MEASURED_TRAFFIC_PER_DOF.get((spec.nx, spec.order, spec.precision))
    dofs = spec.nx * spec.ny * spec.components
    traffic = traffic_dof * dofs / 4 if traffic_dof else None  # bytes per stage launch
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
        "roofline": {"bound": "hbm", "achieved": main_m["achieved_gbs"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": main_m["frac"], "traffic": traffic,
                     "traffic_unit": "bytes per stage launch (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE)",
                     "algorithmic_bytes_per_launch": main_m["bytes_per_dof"] * dofs / 4,
                     "kernel": "stage2d_kernel x4 (one fused RK4 step)",
                     "bytes_per_dof": main_m["bytes_per_dof"],
                     "bytes_per_dof_3reg_scheme": main_m["bytes_per_dof_3reg"],
                     "frac_3reg_equiv": main_m["frac_3reg_equiv"],
                     "stage_ms": main_m["stage_ms"],
                     "copy_ceiling_gbs": copy_gbs},
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
