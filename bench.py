#!/usr/bin/env python3
"""Benchmark of the MI355X MHD step: MHD cell-updates/s and %HBM roofline.

Workload (N=1): 3D ideal-MHD decaying turbulence, 512^3 cells, FP64, periodic,
gamma = 5/3, rho = 1, p = 1, seeded random-phase modes 1 <= |k| <= 4 (rms v =
0.5, rms B = 0.3, B divergence-free).  One step = MISO's ModelBase::update:
CFL reduction, then Integrator::update (4 RK stages + 3 AV sweeps, each
followed by the in-device ghost fill).  Fields are resident in HBM before the
timed region starts.

Prints ONE JSON line.  `roofline` is for the dominant kernel (the fused RK
stage), from HIP events recorded on the library's stream around every launch
of the timed region; `cpu_baseline` times the CPU oracle (a C restatement of
the reference's CPU path) on this host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12          # MI355X HBM3E, B/s
BYTES_STEP_3D = {"f64": 1248, "f32": 624}     # algorithmic bytes per cell-update
# RK stage: stage 1 moves 18 values per cell, stages 2-4 move 27
RK_VALUES = (18, 27, 27, 27)
METRIC = "MHD cell-updates/s & %HBM roofline; RMHD steps/s learned-RT vs plain RT"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512, help="cells per side")
    ap.add_argument("--dtype", choices=("f64", "f32"), default="f64")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target length of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="time a plain device copy of known size (HBM reference for the "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE counters) and exit")
    return ap.parse_args()


def rk_traffic(n, dtype):
    """HBM bytes per rk_stage launch (FETCH_SIZE + WRITE_SIZE, scaled by a
    plain copy's counters) from the committed rocprofv3 summary of this
    workload, or None when there is none."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"mhd_step_turbulence_{n}cubed_{dtype}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cols = line.split("|")
        if line.startswith("| `rk_stage_kernel") and len(cols) >= 10:  # the traffic table
            return float(cols[7]) * 1e6
    return None


def calibrate_copy(nbytes=2 << 30, reps=10):
    """A plain copy kernel of known size: reads and writes `nbytes` per launch."""
    import torch
    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps,
                      "avg_ms": ms, "copy_GBps": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)


def cpu_baseline(target_s: float):
    """The oracle on a bounded sample of the same workload: 3D turbulence,
    FP64, 128^3, as many steps as fit in ~target_s."""
    import numpy as np

    from miso_amd import problems
    from miso_amd.mhd import Grid
    from oracle import oracle as O

    n = 128
    og = O.make_grid((n, n, n), 2)
    g = Grid(n, n, n, 2, device="cpu")
    q = problems.to_numpy(problems.turbulence(g, device="cpu"))
    p = O.Params(cfl_number=0.4)
    O.ghost_fill(og, p, q)
    threads = O.get_threads()
    t0 = time.perf_counter()
    steps = 0
    while True:
        O.run(og, p, q, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= target_s or steps >= 50:
            break
    return {
        "value": n ** 3 * steps / el,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"3D turbulence {n}^3 FP64 (same IC family), {steps} steps in {el:.1f} s, "
                  f"{threads} OpenMP threads",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU (asked for {args.gpus}); the "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    import numpy as np
    import torch

    from miso_amd import hip, mhd, problems

    dev = "cuda:0"
    n = args.n
    dtype = np.float64 if args.dtype == "f64" else np.float32
    grid = mhd.Grid(n, n, n, 2, dtype=dtype, device=dev)
    cfg = mhd.Config({"mhd": {"cfl_number": 0.4}})
    ctx = mhd.ExecContext(0)
    m = mhd.MHD(cfg, grid, ctx, gm=5.0 / 3.0)
    bc = mhd.BoundaryCondition((True, True, True))
    q0 = problems.turbulence(grid, device=dev)
    m.qq.load_ijk(q0)
    del q0
    torch.cuda.synchronize()
    m.integrator.apply_boundary_condition(bc, m.qq)

    for _ in range(args.warmup):
        m.integrator.step(bc, m.qq)
    torch.cuda.synchronize()

    L = hip.lib()
    hip.check(L.misohip_profile_enable(ctx.handle, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dts = []
    for _ in range(args.steps):
        dts.append(m.integrator.step(bc, m.qq))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    import ctypes as C
    nk = len(hip.KIND_NAMES)
    ms = (C.c_double * nk)()
    cnt = (C.c_int * nk)()
    hip.check(L.misohip_profile_read(ctx.handle, ms, cnt, nk))
    hip.check(L.misohip_profile_enable(ctx.handle, 0))

    ncell = grid.n_interior
    es = 8 if args.dtype == "f64" else 4
    value = ncell * args.steps / el
    ms_step = el * 1e3 / args.steps
    kernels = {}
    for k, name in enumerate(hip.KIND_NAMES):
        if cnt[k]:
            kernels[name] = {"launches": cnt[k], "avg_ms": ms[k] / cnt[k]}
    rk = kernels.get("rk_stage")
    rk_bytes = ncell * es * sum(RK_VALUES) * args.steps  # 4 launches per step
    rk_achieved = rk_bytes / (ms[0] * 1e-3) if rk else None
    for name, vals in (("av_x", 18 + 1), ("av_y", 18 + 1), ("av_z", 18 + 1)):
        if name in kernels:
            kernels[name]["achieved_GBps"] = ncell * es * vals / (kernels[name]["avg_ms"] * 1e-3) / 1e9
    if rk:
        kernels["rk_stage"]["achieved_GBps"] = rk_achieved / 1e9
    step_frac = ncell * BYTES_STEP_3D[args.dtype] / (ms_step * 1e-3) / PEAK_HBM

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": f"3D ideal MHD decaying turbulence {n}^3 {args.dtype.upper()}, periodic, "
                        "gamma=5/3, rho=1, p=1, seeded modes 1<=|k|<=4, rms v=0.5, rms B=0.3; "
                        "step = CFL + RK4(4 stages) + 3 AV sweeps + ghost fills",
            "cells": ncell,
            "cfl_number": 0.4,
        },
        "hbm_roofline_frac_step": step_frac,
        "roofline": {
            "bound": "hbm",
            "kernel": "rk_stage_kernel (fused 9-field RK stage)",
            "achieved": rk_achieved / 1e9 if rk else None,
            "peak": PEAK_HBM / 1e9,
            "unit": "GB/s",
            "frac": (rk_achieved / PEAK_HBM) if rk else None,
            "traffic": rk_traffic(n, args.dtype),
            "traffic_source": f"profiles/mhd_step_turbulence_{n}cubed_{args.dtype}.md",
        },
        "kernels": kernels,
        "dt_last": dts[-1],
    }
    ctx.close()
    del m
    torch.cuda.empty_cache()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
