#!/usr/bin/env python3
"""Benchmark of the MI355X MHD step: MHD cell-updates/s and %HBM roofline.

Workload (N=1): 3D ideal-MHD decaying turbulence, 512^3 cells, FP64, periodic,
gamma = 5/3, rho = 1, p = 1, seeded random-phase modes 1 <= |k| <= 4 (rms v =
0.5, rms B = 0.3, B divergence-free).  One step = MISO's ModelBase::update:
CFL reduction, then Integrator::update (4 RK stages + 3 AV sweeps, each
followed by the in-device ghost fill).  Fields are resident in HBM before the
timed region starts.

Then the radiative-MHD leg (`rmhd`): steps/s of the stratified radiative box
with a plain GPU RT solve on every step (256^3 and 384^3) and with the
learned RT (384^3: the per-column network on every step, a plain solve every
16 steps and on flagged columns), the network trained in the same run on
columns of the GPU RT solver, and its Q_rad error against converged solves.

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
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--tiles", default=None,
                    help="bx,by,chunk,variant[,zchunk[,fuse_xy[,xy_chunk]]]: set the 3D tile "
                         "shapes instead of tuning")
    ap.add_argument("--no-rmhd", action="store_true", help="skip the radiative-MHD leg")
    ap.add_argument("--rmhd-only", action="store_true", help="run only the radiative-MHD leg")
    ap.add_argument("--rt-only", action="store_true",
                    help="run only the plain-RT part of the radiative leg (args.rmhd_n^3)")
    ap.add_argument("--rmhd-n", type=int, default=384, help="learned-RT box, cells per side")
    ap.add_argument("--rmhd-plain-n", type=int, default=256, help="plain-RT box, cells per side")
    ap.add_argument("--rmhd-plain-steps", type=int, default=8)
    ap.add_argument("--rmhd-cycles", type=int, default=3,
                    help="timed learned-RT cycles of rt_every steps")
    ap.add_argument("--rmhd-rt-every", type=int, default=16)
    ap.add_argument("--rmhd-flag-tol", type=float, default=None,
                    help="columns whose network misses the solve by more than this take the "
                         "solved values (default: no flags)")
    ap.add_argument("--rmhd-no-correct", action="store_true",
                    help="do not add the correction the last plain solve leaves")
    ap.add_argument("--rmhd-train-columns", type=int, default=400000)
    ap.add_argument("--rmhd-train-seeds", type=int, default=6,
                    help="boxes (other seeds) the training columns come from")
    ap.add_argument("--rmhd-epochs", type=int, default=400)
    ap.add_argument("--rmhd-slab-tol", type=float, default=None,
                    help="tolerance of the bottom-slab solves (default: the RT tolerance)")
    ap.add_argument("--rmhd-slab-iters", type=int, default=2,
                    help="iterations of a bottom-slab solve (at most; the slab starts from its "
                         "last solve, or from the last full solve after one)")
    ap.add_argument("--rmhd-solve-iters", type=int, default=None,
                    help="most iterations of the learned run's plain solve every rt_every "
                         "steps (default: the RT limit)")
    ap.add_argument("--rmhd-slab-share", type=float, default=0.99,
                    help="share of the network's residual at a solve whose bottom layers are "
                         "solved plainly on every step (<= 0: no slab)")
    ap.add_argument("--rmhd-features", type=int, default=4, choices=(3, 4),
                    help="column inputs: 3 = (ln ro, ln T, B^2), 4 = plus the horizontal "
                         "Laplacian of ln T")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--ensemble-only", action="store_true")
    ap.add_argument("--ens-runs", type=int, default=64)
    ap.add_argument("--ens-seed", type=int, default=2026)
    ap.add_argument("--ens-graph-steps", type=int, default=10)
    ap.add_argument("--ens-replays", type=int, default=5)
    ap.add_argument("--ens-chunk", type=int, default=0,
                    help="walk length of every run's y sweep (default: tuned per run size)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the Orszag-Tang accuracy leg")
    ap.add_argument("--accuracy-only", action="store_true")
    ap.add_argument("--no-f32", action="store_true", help="skip the FP32 timing of the 3D step")
    ap.add_argument("--trace-markers", action="store_true",
                    help="launch a one-block int16 fill kernel right before and right after "
                         "each timed region of the radiative leg (markers 1-2: plain RT at "
                         "--rmhd-plain-n, 3-4: plain RT at --rmhd-n, 5-6: learned RT), so a "
                         "kernel trace can be cut to those steps (--trace-window)")
    ap.add_argument("--trace-window", nargs=3, metavar=("CSV", "FIRST", "STEPS"), default=None,
                    help="read a rocprofv3 kernel-trace CSV, take the kernels between marker "
                         "FIRST and marker FIRST+1 (counted from 1), print their time per step "
                         "over STEPS steps as a markdown table, and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="time a plain device copy of known size (HBM reference for the "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE counters) and exit")
    return ap.parse_args()


def rk_kernel_name(variant, dtype="f64"):
    """Name (as the profiles list it) of the RK stage kernel a 3D run
    launches for a tile variant; for the z-march variants the prefix shared
    by the first-stage and later-stage instances."""
    t = "double" if dtype == "f64" else "float"
    return {0: f"rk_stage_kernel<{t}, true, true, true, 0>",
            1: f"rk_stage_kernel<{t}, true, true, true, 1>",
            2: f"rk_stage_kernel_w4<{t}, true, true, true>", 3: f"rk_stage_kernel_tile<{t}>",
            4: f"rk_stage_kernel_march<{t}, 4", 5: f"rk_stage_kernel_march<{t}, 8",
            6: f"rk_stage_kernel_march<{t}, {12 if dtype == 'f32' else 8}"}[variant]


def rk_traffic(n, dtype, variant=0):
    """HBM bytes per launch of the RK stage kernel of this variant
    (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy's counters) from the
    committed rocprofv3 summary of this workload, or None when there is
    none."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"mhd_step_turbulence_{n}cubed_{dtype}.md")
    if not os.path.exists(path):
        return None
    name = "| `" + rk_kernel_name(variant, dtype)
    tot, cnt = 0.0, 0
    in_traffic = False
    for line in open(path):
        if line.startswith("| kernel |"):  # a table header: is it the traffic table?
            in_traffic = "traffic MB" in line
            continue
        cols = line.split("|")
        if in_traffic and line.startswith(name) and len(cols) >= 11:
            # instances of one variant (the z-march first stage has its own)
            # are averaged over their dispatches
            tot += float(cols[7]) * 1e6 * int(cols[2])
            cnt += int(cols[2])
    return tot / cnt if cnt else None


def calibrate_copy(nbytes=2 << 30, reps=10):
    """Plain streaming kernels of known size: torch's device copy, and the
    library's hand-written copy, read and write kernels (misohip_stream_bench:
    16-byte accesses, nontemporal stores) on 4 GiB buffers -- the HBM rate the
    MHD kernels are compared against.  The torch copy of `nbytes` is also the
    kernel the counter calibration profiles."""
    import ctypes as C

    import torch

    from miso_amd import hip
    from miso_amd.mhd import ExecContext
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
    out = {"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps,
           "avg_ms": ms, "copy_GBps": 2 * nbytes / (ms * 1e-3) / 1e9}
    del a, b
    big = 4 << 30
    x = torch.ones(big // 8, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    ctx = ExecContext(0)
    kinds = {"copy": (0, 2), "read": (1, 1), "write": (2, 1)}
    for name, (kind, passes) in kinds.items():
        args = (ctx.handle, kind, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), big)
        hip.call("misohip_stream_bench", *args)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            e0.record()
            hip.call("misohip_stream_bench", *args)
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        out[f"hand_{name}_GBps"] = passes * big / (best * 1e-3) / 1e9
    ctx.close()
    print(json.dumps(out), flush=True)


def cpu_baseline(target_s: float, q_full=None, n_full: int = 0):
    """The oracle on the benchmark's own workload: with q_full (the GPU run's
    initial state, MISO order, ghosts filled) one step -- two when the first
    took less than target_s / 2 -- of the n_full^3 FP64 turbulence on every
    host thread; without it a 128^3 sample of the same initial-condition
    family for ~target_s.  A 128^3 step on one thread gives the per-thread
    rate."""
    import numpy as np

    from miso_amd import problems
    from miso_amd.mhd import Grid
    from oracle import oracle as O

    p = O.Params(cfl_number=0.4)
    threads = O.get_threads()
    if q_full is not None:
        n = n_full
        og = O.make_grid((n, n, n), 2)
        q = q_full
        t0 = time.perf_counter()
        steps = 0
        while True:
            O.run(og, p, q, 1)
            steps += 1
            el = time.perf_counter() - t0
            if steps >= 2 or el >= 0.5 * target_s:
                break
        del og
    n1 = 128
    og1 = O.make_grid((n1, n1, n1), 2)
    g1 = Grid(n1, n1, n1, 2, device="cpu")
    q1 = problems.to_numpy(problems.turbulence(g1, device="cpu"))
    O.ghost_fill(og1, p, q1)
    if q_full is None:
        n, og, q = n1, og1, q1
        t0 = time.perf_counter()
        steps = 0
        while True:
            O.run(og, p, q, 1)
            steps += 1
            el = time.perf_counter() - t0
            if el >= target_s or steps >= 50:
                break
    # one more 128^3 step on one thread (the per-thread rate)
    O.set_threads(1)
    t1 = time.perf_counter()
    O.run(og1, p, q1, 1)
    el1 = time.perf_counter() - t1
    O.set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                         "")
    except OSError:
        pass
    return {
        "value": n ** 3 * steps / el,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"3D turbulence {n}^3 FP64, the benchmark's initial state, {steps} "
                   f"step(s) in {el:.1f} s, {threads} OpenMP threads" if q_full is not None else
                   f"3D turbulence {n}^3 FP64 (same IC family), {steps} steps in {el:.1f} s, "
                   f"{threads} OpenMP threads"),
        "single_thread_value": n1 ** 3 / el1,
        "single_thread_sample": f"one {n1}^3 step on one thread",
        "host": {"cpu": model, "nproc": os.cpu_count()},
    }


BYTES_STEP_2D = {"f64": 1096, "f32": 548}     # algorithmic bytes per 2D cell-update


def _div_b(q, dx, dy, dz):
    """max |div B| dx / max |B| over the interior (4th-order central
    differences), q in MISO order (9, It, Jt, Kt) with margin 2 (0 on a
    size-1 axis)."""
    import numpy as np

    def d(f, axis, h):
        if f.shape[axis] == 1:
            return np.zeros_like(f)
        s = [slice(None)] * 3
        def sh(o):
            s2 = list(s)
            s2[axis] = slice(2 + o, f.shape[axis] - 2 + o)
            return f[tuple(s2)]
        return (-sh(2) + 8 * sh(1) - 8 * sh(-1) + sh(-2)) / (12 * h)
    bx, by, bz = (q[4].astype(np.float64), q[5].astype(np.float64), q[6].astype(np.float64))
    parts = []
    for f, ax, h in ((bx, 0, dx), (by, 1, dy), (bz, 2, dz)):
        g = d(f, ax, h)
        # trim the other axes to their interiors
        sl = [slice(2, -2) if (a != ax and f.shape[a] > 1) else slice(None) for a in range(3)]
        parts.append(g[tuple(sl)])
    div = parts[0] + parts[1] + parts[2]
    # |B| over the interior (edge and corner ghost cells are not part of
    # the state: the oracle leaves them unset)
    inner = tuple(slice(2, -2) if bx.shape[a] > 1 else slice(None) for a in range(3))
    bmag = np.sqrt(bx[inner] ** 2 + by[inner] ** 2 + bz[inner] ** 2).max()
    return float(np.abs(div).max() * min(dx, dy, dz) / bmag)


def f32_leg(args):
    """The same 3D turbulence step in FP32 (tuned tiles), cell-updates/s."""
    import numpy as np
    import torch

    from miso_amd import mhd, problems
    n = args.n
    grid = mhd.Grid(n, n, n, 2, dtype=np.float32, device="cuda:0")
    ctx = mhd.ExecContext(0)
    m = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), grid, ctx, gm=5.0 / 3.0)
    bc = mhd.BoundaryCondition((True, True, True))
    m.qq.load_ijk(problems.turbulence(grid, device="cuda:0"))
    m.integrator.apply_boundary_condition(bc, m.qq)
    tiles = None
    if not args.no_tune:
        from miso_amd.autotune import summary, tune_single
        tiles = summary(tune_single(ctx, m, bc, steps=1))
        tiles.pop("times_ms", None)
    for _ in range(args.warmup):
        m.integrator.step(bc, m.qq)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.integrator.run(bc, m.qq, args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    v = grid.n_interior * args.steps / el
    out = {"cell_updates_per_s": v, "ms_per_step": 1e3 * el / args.steps,
           "hbm_roofline_frac_step": v * BYTES_STEP_3D["f32"] / PEAK_HBM, "tiles": tiles}
    print(f"[f32] {out}", file=sys.stderr, flush=True)
    ctx.close()
    del m
    torch.cuda.empty_cache()
    return out


def accuracy_leg(args):
    """Orszag-Tang 256^2 (FP64 and FP32 GPU runs of 200 steps at CFL 0.4)
    against the FP64 oracle: parity of the FP64 run, per-field error of the
    FP32 run, div B of both, GPU and CPU times of the same 200 steps."""
    import numpy as np
    import torch

    from miso_amd import mhd, problems
    from oracle import oracle as O

    n, nsteps = 256, 200
    ctx = mhd.ExecContext(0)
    cfg = mhd.Config({"mhd": {"cfl_number": 0.4}})
    bc = mhd.BoundaryCondition((True, True, True))
    og = O.make_grid((n, n, 1), 2)
    g64 = mhd.Grid(n, n, 1, 2, device="cuda:0")
    q0 = problems.to_numpy(problems.orszag_tang(g64))
    ref = q0.copy()
    p = O.Params(cfl_number=0.4)
    O.ghost_fill(og, p, ref)
    t0 = time.perf_counter()
    ref_dts = O.run(og, p, ref, nsteps)
    cpu_s = time.perf_counter() - t0
    out = {"workload": f"2D Orszag-Tang {n}^2, gamma 5/3, periodic, CFL 0.4, {nsteps} steps",
           "cpu_oracle_seconds": cpu_s, "cpu_threads": O.get_threads()}
    sl = (slice(None),) + g64.interior()
    names = ("ro", "vx", "vy", "vz", "bx", "by", "bz", "ei", "ph")
    for dt_name, dt in (("f64", np.float64), ("f32", np.float32)):
        g = g64 if dt == np.float64 else mhd.Grid(n, n, 1, 2, dtype=np.float32, device="cuda:0")
        m = mhd.MHD(cfg, g, ctx, 5.0 / 3.0)
        m.apply_initial_condition(q0.astype(dt), bc)
        div0 = _div_b(m.qq.to_host(ctx), 1.0 / n, 1.0 / n, 1.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dts = [m.integrator.step(bc, m.qq) for _ in range(nsteps)]
        torch.cuda.synchronize()
        gpu_s = time.perf_counter() - t0
        q = m.qq.to_host(ctx).astype(np.float64)
        err = {}
        for f, nm in enumerate(names):
            a, b = q[sl][f], ref[sl][f]
            scale = max(np.abs(b).max(), 1e-300)
            err[nm] = {"rel_linf": float(np.abs(a - b).max() / scale),
                       "rel_l2": float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))}
        out[dt_name] = {"gpu_seconds": gpu_s, "gpu_cell_updates_per_s": n * n * nsteps / gpu_s,
                        "dt_equal_to_oracle_every_step": bool(np.array_equal(np.array(dts),
                                                                             ref_dts)),
                        "error_vs_fp64_oracle": err,
                        "max_rel_linf": max(e["rel_linf"] for e in err.values()),
                        "div_b_rel_start": div0,
                        "div_b_rel_end": _div_b(q, 1.0 / n, 1.0 / n, 1.0)}
        if dt == np.float64:
            # the same 200 steps as a one-run batch replayed from a hipGraph
            # (dt stays on the device: no host round trip per step)
            from miso_amd.ensemble import Ensemble
            q0t = torch.from_numpy(q0)
            ens = Ensemble(ctx, [(n, n, 1)], lambda _m, _g: q0t, bc=bc, dtype=np.float64,
                           cfl_number=0.4, gm=5.0 / 3.0,
                           bounds=((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)))
            ens.capture(50)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ens.step(nsteps)
            torch.cuda.synchronize()
            graph_s = time.perf_counter() - t0
            qg = ens.runs[0].qq.to_host(ctx)
            out[dt_name]["graph"] = {
                "gpu_seconds": graph_s, "gpu_cell_updates_per_s": n * n * nsteps / graph_s,
                "same_bits_as_step_by_step": bool(np.array_equal(qg, m.qq.to_host(ctx))),
                "last_dt_equal": bool(float(ens.state()[0][0]) == float(dts[-1]))}
            ens.close()
        del m
    print(f"[accuracy] fp64 max rel linf {out['f64']['max_rel_linf']:.2e}, "
          f"fp32 {out['f32']['max_rel_linf']:.2e}", file=sys.stderr, flush=True)
    out["turbulence_3d"] = accuracy_3d(ctx, args)
    ctx.close()
    return out


def accuracy_3d(ctx, args, n=64, nsteps=50):
    """3D decaying turbulence n^3 (the benchmark's initial-condition family),
    nsteps of misohip_run at CFL 0.4 in FP64 and in FP32 against the FP64
    oracle: parity and dts of the FP64 run, per-field error of the FP32 run,
    div B of both at the start and the end."""
    import numpy as np
    import torch

    from miso_amd import mhd, problems
    from oracle import oracle as O

    cfg = mhd.Config({"mhd": {"cfl_number": 0.4}})
    bc = mhd.BoundaryCondition((True, True, True))
    og = O.make_grid((n, n, n), 2)
    g64 = mhd.Grid(n, n, n, 2, device="cuda:0")
    q0 = problems.to_numpy(problems.turbulence(g64))
    ref = q0.copy()
    p = O.Params(cfl_number=0.4)
    O.ghost_fill(og, p, ref)
    t0 = time.perf_counter()
    ref_dts = O.run(og, p, ref, nsteps)
    cpu_s = time.perf_counter() - t0
    h = 1.0 / n
    # the oracle's run ends without a ghost fill (misohip_run ends with one):
    # div B reads face ghosts, so it is taken on a filled copy
    ref_filled = ref.copy()
    O.ghost_fill(og, p, ref_filled)
    out = {"workload": f"3D turbulence {n}^3, gamma 5/3, periodic, CFL 0.4, {nsteps} steps "
                       "(misohip_run)",
           "cpu_oracle_seconds": cpu_s, "div_b_rel_oracle_end": _div_b(ref_filled, h, h, h)}
    sl = (slice(None),) + g64.interior()
    names = ("ro", "vx", "vy", "vz", "bx", "by", "bz", "ei", "ph")
    for dt_name, dt in (("f64", np.float64), ("f32", np.float32)):
        g = g64 if dt == np.float64 else mhd.Grid(n, n, n, 2, dtype=np.float32, device="cuda:0")
        m = mhd.MHD(cfg, g, ctx, 5.0 / 3.0)
        m.apply_initial_condition(q0.astype(dt), bc)
        div0 = _div_b(m.qq.to_host(ctx), h, h, h)
        dts = m.integrator.run(bc, m.qq, nsteps)
        torch.cuda.synchronize()
        q = m.qq.to_host(ctx).astype(np.float64)
        err = {}
        for f, nm in enumerate(names):
            a, b = q[sl][f], ref[sl][f]
            scale = max(np.abs(b).max(), 1e-300)
            err[nm] = {"rel_linf": float(np.abs(a - b).max() / scale),
                       "rel_l2": float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))}
        r = {"error_vs_fp64_oracle": err,
             "max_rel_linf": max(e["rel_linf"] for e in err.values()),
             "dt_equal_to_oracle_every_step": bool(np.array_equal(np.array(dts), ref_dts)),
             "dt_last_rel_diff": float(abs(dts[-1] - ref_dts[-1]) / ref_dts[-1]),
             "div_b_rel_start": div0, "div_b_rel_end": _div_b(q, h, h, h)}
        if dt == np.float64:
            r["bitwise_equal_to_oracle"] = bool(np.array_equal(q[sl], ref[sl]))
        out[dt_name] = r
        del m
    print(f"[accuracy 3d] fp64 max rel linf {out['f64']['max_rel_linf']:.2e}, fp32 "
          f"{out['f32']['max_rel_linf']:.2e}", file=sys.stderr, flush=True)
    return out


def ensemble_leg(args):
    """Batched ensemble: args.ens_runs independent 2D Kelvin-Helmholtz runs,
    FP32, sizes drawn by seed from {128^2, 256^2, 512^2, 1024^2} (density
    ratio 2, dv = 1, tanh width 0.02, seeded Bx in [0, 0.2], 1e-3 seeded vy
    noise, gamma 1.4, p 2.5, x periodic, y mirror), advanced together by
    replays of a hipGraph of args.ens_graph_steps batched steps; each run
    keeps its own dt on the device.  Beside it, the same runs stepped one
    after another through misohip_step (the unbatched path)."""
    import numpy as np
    import torch

    from miso_amd import problems
    from miso_amd.ensemble import KH_BC, Ensemble, kh_sizes
    from miso_amd.mhd import ExecContext

    ctx = ExecContext(0)
    sizes = kh_sizes(args.ens_runs, args.ens_seed)

    def init(m, g):
        return problems.kelvin_helmholtz(g, seed=args.ens_seed + m, device="cpu")

    t0 = time.perf_counter()
    ens = Ensemble(ctx, [(n, n, 1) for n in sizes], init, dtype=np.float32)
    setup_s = time.perf_counter() - t0
    tuned = None
    if args.ens_chunk > 0:
        ens.set_chunks([args.ens_chunk] * ens.n)
        tuned = {"chunks": {"all": args.ens_chunk}}
    elif not args.no_tune:
        from miso_amd.autotune import summary, tune_ensemble
        tuned = summary(tune_ensemble(
            ctx, ens, lambda s: Ensemble(ctx, [s, s], init, dtype=np.float32)))
        tuned.pop("times_ms", None)
    k = args.ens_graph_steps
    ens.capture(k)
    ens.step(k)  # warm-up replay
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(k * args.ens_replays)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    steps = k * args.ens_replays
    cells = ens.cells
    value = cells * steps / el
    dt, tm, st, fb = ens.state()
    # unbatched: every run through misohip_step, one after another
    t0 = time.perf_counter()
    seq_steps = 2
    for _ in range(seq_steps):
        for r in ens.runs:
            r.integrator.step(KH_BC, r.qq)
    torch.cuda.synchronize()
    seq = time.perf_counter() - t0
    out = {
        "workload": f"{len(sizes)} independent 2D Kelvin-Helmholtz runs FP32, sizes drawn by seed "
                    f"{args.ens_seed} from 128^2..1024^2, one hipGraph of {k} batched steps per "
                    "replay, per-run dt on the device",
        "runs": len(sizes),
        "size_counts": {str(n): sizes.count(n) for n in sorted(set(sizes))},
        "cells": cells,
        "steps": steps,
        "ms_per_step": 1e3 * el / steps,
        "cell_updates_per_s": value,
        "hbm_roofline_frac": value * BYTES_STEP_2D["f32"] / PEAK_HBM,
        "unbatched_cell_updates_per_s": cells * seq_steps / seq,
        "speedup_vs_unbatched": value / (cells * seq_steps / seq),
        "exp_fallback": fb,
        "tiles": tuned,
        "dt_range": [float(dt.min()), float(dt.max())],
        "setup_seconds": setup_s,
    }
    print(f"[ensemble] {out}", file=sys.stderr, flush=True)
    ens.close()
    del ens
    ctx.close()
    torch.cuda.empty_cache()
    return out


TRACE_MARKER = "FillFunctor<short>"


def trace_marker():
    """A one-block int16 fill on the current stream, after the device is
    idle: no other kernel of the benchmark fills int16, so it marks a point
    in a kernel trace."""
    import torch

    torch.cuda.synchronize()
    torch.full((64,), 1, dtype=torch.int16, device="cuda:0")
    torch.cuda.synchronize()


def trace_window(path, first, steps):
    """Per-kernel time per step between trace markers first and first + 1
    of a rocprofv3 kernel-trace CSV: launches per step, ms per step and
    share of the kernel time, then the window's wall time per step (its
    markers' ends and starts) and its summed kernel time per step (kernels
    on two streams can overlap, so the sum can exceed the wall time)."""
    import csv
    import re

    names = {}  # kernel names interned: a training run traces millions of dispatches
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            n = names.setdefault(r["Kernel_Name"], r["Kernel_Name"])
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), n))
    rows.sort(key=lambda t: t[0])
    marks = [i for i, t in enumerate(rows) if TRACE_MARKER in t[2]]
    if len(marks) < first + 1:
        sys.exit(f"trace_window: {len(marks)} markers in {path}, need {first + 1}")
    a, b = marks[first - 1], marks[first]
    wall = rows[b][0] - rows[a][1]
    per = {}
    busy = 0
    for s, e, n in rows[a + 1:b]:
        name = re.sub(r"\(.*$", "", re.sub(r"^void ", "", n)).replace("misohip::", "")
        c = per.setdefault(name, [0, 0])
        c[0] += 1
        c[1] += e - s
        busy += e - s
    print("| kernel | launches per step | ms per step | share of kernel time |")
    print("|---|---|---|---|")
    for name, (c, t) in sorted(per.items(), key=lambda kv: -kv[1][1]):
        if t / 1e6 / steps >= 0.005:
            print(f"| `{name[:90]}` | {c / steps:.1f} | {t / 1e6 / steps:.3f} | "
                  f"{100.0 * t / busy:.1f} % |")
    print(f"\nwindow: {wall / 1e6 / steps:.3f} ms per step wall, {busy / 1e6 / steps:.3f} ms "
          f"per step of kernel time, {steps:g} steps")


def rmhd_leg(args):
    """Radiative MHD on the stratified box: plain GPU RT on every step
    (args.rmhd_plain_n^3 and args.rmhd_n^3) against the learned RT
    (args.rmhd_n^3: the column network on every step, a plain solve every
    rt_every steps and on flagged columns).  The network is trained here on
    columns the GPU RT solver produces for boxes with other seeds; its
    Q_rad is checked against converged RT solves of the same states."""
    import numpy as np
    import torch

    from miso_amd.mhd import ExecContext, Grid
    from miso_amd.rmhd import RadiativeMHD, rel_l2
    from miso_amd.surrogate import ColumnSurrogate, collect_columns

    dev = "cuda:0"
    ctx = ExecContext(0)

    def log(s):
        print(f"[rmhd] {s}", file=sys.stderr, flush=True)

    def make_box(n, seed):
        g = Grid(n, n, n, bounds=((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), device=dev)
        b = RadiativeMHD(g, ctx)
        b.initialize(seed)
        if not args.no_tune:  # the MHD step's tiles for this size (cached after the first box)
            from miso_amd.autotune import tune_single
            tune_single(ctx, b.model, b.bc, steps=1)
        return b

    def time_plain(n, steps, seed):
        box = make_box(n, seed)
        for _ in range(2):
            box.step_plain()
        torch.cuda.synchronize()
        i0 = len(box.rt_iters)
        evs = []
        if args.trace_markers:
            trace_marker()
        t0 = time.perf_counter()
        for _ in range(steps):  # box.step_plain(), with events between its two parts
            e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e[0].record()
            q = box.solve_rt()
            e[1].record()
            box._advance(q)
            e[2].record()
            evs.append(e)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        if args.trace_markers:
            trace_marker()
        rt_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
        mhd_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
        # a learned RT can at best remove the solve: (MHD + RT) / MHD
        r = {"cells": n ** 3, "steps": steps, "ms_per_step": 1e3 * el / steps,
             "steps_per_s": steps / el, "rt_iterations_mean": float(np.mean(box.rt_iters[i0:])),
             "rt_solve_ms": rt_ms, "mhd_step_ms": mhd_ms,
             "learned_speedup_bound": (rt_ms + mhd_ms) / mhd_ms}
        log(f"plain RT {n}^3: {r}")
        del box
        torch.cuda.empty_cache()
        return r

    out = {"workload": "stratified radiative MHD box, FP64 fields, T 2 -> 0.5, scale height "
                       "0.1 box, kappa = kappa0 ro T^3 (tau = 1 at mid-height), 24 rays, "
                       "1e-2 seeded v and B perturbations, x/y periodic, fixed z ghosts"}
    n = args.rmhd_n
    if args.rt_only:
        out["plain_rt"] = time_plain(n, args.rmhd_plain_steps, seed=7)
        ctx.close()
        return out
    out["plain_rt_small"] = time_plain(args.rmhd_plain_n, args.rmhd_plain_steps, seed=7)
    out["plain_rt"] = time_plain(n, args.rmhd_plain_steps, seed=7)
    every = args.rmhd_rt_every
    nsteps = every * args.rmhd_cycles
    # training data from other seeds, over the same span of steps
    t0 = time.perf_counter()
    data_box = make_box(n, 101)
    F, T = collect_columns(data_box, seeds=tuple(range(101, 101 + args.rmhd_train_seeds)),
                           steps=nsteps + every,
                           sample_every=4, max_columns=args.rmhd_train_columns, log=log,
                           features=args.rmhd_features)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    del data_box
    torch.cuda.empty_cache()
    log(f"training data: {F.shape[0]} columns in {gen_s:.1f} s")
    sur = ColumnSurrogate(ctx, n, hidden=256, hidden_layers=3, max_batch=n * n, seed=0,
                          features=args.rmhd_features)
    t0 = time.perf_counter()
    mse, val = sur.fit(F, T, epochs=args.rmhd_epochs, batch=1024, lr=1e-3, log_every=5, log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    ncols = int(F.shape[0])
    del F, T
    torch.cuda.empty_cache()
    log(f"trained: mse {mse:.3e}, validation rel L2 {val:.3e}, {train_s:.1f} s")
    slab_share = args.rmhd_slab_share if args.rmhd_slab_share > 0 else None
    # timed learned run: one untimed cycle, then args.rmhd_cycles cycles
    box = make_box(n, 7)
    box.attach_surrogate(sur, rt_every=every, flag_tol=args.rmhd_flag_tol,
                         correct=not args.rmhd_no_correct, slab_share=slab_share,
                         slab_tolerance=args.rmhd_slab_tol,
                         slab_max_iters=args.rmhd_slab_iters,
                         solve_max_iters=args.rmhd_solve_iters)
    for _ in range(every):
        box.step_learned()
    if args.trace_markers:
        trace_marker()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        box.step_learned()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if args.trace_markers:
        trace_marker()
    learned = {"cells": n ** 3, "steps": nsteps, "rt_every": every,
               "ms_per_step": 1e3 * el / nsteps, "steps_per_s": nsteps / el,
               "flagged_column_fraction": float(box.flags.float().mean()),
               "bottom_slab": None if box.slab is None else {
                   "kept_layers": box.slab.layers, "buffer_layers": box.slab.buffer,
                   "residual_share": slab_share, "tolerance": box.slab_kw["tolerance"],
                   "max_iters": box.slab_kw["max_iters"],
                   "rt_iterations_mean": float(np.mean(box.slab.iters)) if box.slab.iters
                   else None}}
    del box
    torch.cuda.empty_cache()
    # accuracy: the same run again, Q_rad checked against converged solves
    box = make_box(n, 7)
    box.attach_surrogate(sur, rt_every=every, flag_tol=args.rmhd_flag_tol,
                         correct=not args.rmhd_no_correct, slab_share=slab_share,
                         slab_tolerance=args.rmhd_slab_tol,
                         slab_max_iters=args.rmhd_slab_iters,
                         solve_max_iters=args.rmhd_solve_iters)
    errs = {"learned": [], "network_only": [], "network_corrected": [], "stale_rt": []}
    by_age = {}
    tmp = torch.empty_like(box.q_rt)

    def probe(q, solved):
        s = box.steps_done
        if solved or s % 3:
            return
        truth = box.reference_qrad()
        errs["learned"].append(rel_l2(q, truth, box.grid))
        by_age.setdefault(s % every, []).append(errs["learned"][-1])
        errs["stale_rt"].append(rel_l2(box.q_rt, truth, box.grid))
        box.learned_qrad(out=tmp, correct=False, use_flags=False)
        errs["network_only"].append(rel_l2(tmp, truth, box.grid))
        if box.corr is not None:
            box.learned_qrad(out=tmp, correct=True, use_flags=False)
            errs["network_corrected"].append(rel_l2(tmp, truth, box.grid))

    for s in range(nsteps + every):
        box.step_learned(probe)
        if s % 8 == 0:
            log(f"accuracy pass step {s}")
    for k, v in errs.items():
        if v:
            learned[f"qrad_rel_l2_{k}"] = {"mean": float(np.mean(v)), "max": float(np.max(v)),
                                           "checks": len(v)}
    learned["qrad_rel_l2_learned_by_steps_since_solve"] = {
        str(a): float(np.mean(v)) for a, v in sorted(by_age.items())}
    learned["training"] = {"columns": ncols, "epochs": args.rmhd_epochs, "train_mse": mse,
                           "seeds": args.rmhd_train_seeds,
                           "features": args.rmhd_features,
                           "validation_rel_l2": val, "data_seconds": gen_s,
                           "train_seconds": train_s}
    log(f"learned RT {n}^3: {learned}")
    out["learned_rt"] = learned
    out["speedup_learned_vs_plain"] = learned["steps_per_s"] / out["plain_rt"]["steps_per_s"]
    sur.close()
    del box
    ctx.close()
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on exactly one GPU (asked for {args.gpus}); the "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.trace_window:
        trace_window(args.trace_window[0], int(args.trace_window[1]),
                     float(args.trace_window[2]))
        return
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.rt_only:
        print(json.dumps({"rt": rmhd_leg(args)}), flush=True)
        return
    if args.rmhd_only:
        print(json.dumps({"rmhd": rmhd_leg(args)}), flush=True)
        return
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_leg(args)}), flush=True)
        return
    if args.accuracy_only:
        print(json.dumps({"accuracy": accuracy_leg(args)}), flush=True)
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
    # the initial state in MISO order for the CPU baseline (FP64 runs)
    q_host = m.qq.to_host(ctx) if (not args.no_cpu and args.dtype == "f64") else None
    tiles = None
    if args.tiles:
        from miso_amd.autotune import set_tiles
        bx, by, ch, var, *rest = (int(x) for x in args.tiles.split(","))
        zc = rest[0] if rest else 32
        fz = rest[1] if len(rest) > 1 else 0
        xc = rest[2] if len(rest) > 2 else 64
        set_tiles(ctx, bx, by, ch, var, zc, fz, xc)
        tiles = {"rk_pencil": [bx, by], "sweep_chunk": ch, "rk_variant": var, "rk_zchunk": zc,
                 "av_fuse_xy": fz, "av_xy_chunk": xc, "given": True}
    elif not args.no_tune:
        from miso_amd.autotune import summary, tune_single
        tiles = summary(tune_single(ctx, m, bc, steps=1))
        print(f"[tune] {tiles}", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        m.integrator.step(bc, m.qq)
    torch.cuda.synchronize()

    L = hip.lib()
    hip.check(L.misohip_profile_enable(ctx.handle, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dts = m.integrator.run(bc, m.qq, args.steps)
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
    if tiles and tiles.get("av_fuse_xy", 0) >= 3 and "av_x" in kernels and "av_z" not in kernels:
        # the z walk overlapped with the x-y walk: one record for both
        kernels["av_xy_and_z_overlapped"] = kernels.pop("av_x")
    for name, vals in (("av_x", 18 + 1), ("av_y", 18 + 1), ("av_z", 18 + 1),
                       ("av_xy_and_z_overlapped", 2 * (18 + 1))):
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
                        "step = CFL + RK4(4 stages) + 3 AV sweeps + ghost fills; K steps in one "
                        "misohip_run call (each later step's CFL reduced in the previous "
                        "step's last sweep)",
            "cells": ncell,
            "cfl_number": 0.4,
        },
        "hbm_roofline_frac_step": step_frac,
        "roofline": {
            "bound": "hbm",
            "kernel": rk_kernel_name((tiles or {}).get("rk_variant", 0), args.dtype) +
                      (", *>" if (tiles or {}).get("rk_variant", 0) in (4, 5) else "") +
                      " (fused 9-field RK stage)",
            "achieved": rk_achieved / 1e9 if rk else None,
            "peak": PEAK_HBM / 1e9,
            "unit": "GB/s",
            "frac": (rk_achieved / PEAK_HBM) if rk else None,
            "traffic": rk_traffic(n, args.dtype, (tiles or {}).get("rk_variant", 0)),
            "traffic_source": f"profiles/mhd_step_turbulence_{n}cubed_{args.dtype}.md",
        },
        "kernels": kernels,
        "tiles": tiles,
        "dt_last": dts[-1],
    }
    ctx.close()
    del m
    torch.cuda.empty_cache()
    if not args.no_f32 and args.dtype == "f64":
        out["f32_step"] = f32_leg(args)
    if not args.no_accuracy:
        out["accuracy"] = accuracy_leg(args)
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(args)
    if not args.no_rmhd:
        out["rmhd"] = rmhd_leg(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, q_host, n)
        del q_host
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
