"""Tile autotuner (no counterpart in the reference).

Picks, per device, kernel and problem size, the tile shapes the HIP kernels
expose, by timing the candidates on the real state:

- single runs (`tune_single`): the RK stage's pencil shape
  (`rk_pencil_bx` x `rk_pencil_by` blocks of 64 x 4 cells walked along z),
  its variant (`rk_variant`, 3D: load phases, register cap, LDS-staged x-y
  neighbourhood, z-march with its planes per block `rk_zchunk`), the walk
  length of the y/z AV sweeps (`sweep_chunk`) and, in 3D, whether the x and
  y sweeps run as one walk (`av_fuse_xy`, with its own `av_xy_chunk`),
  timed with the
  library's own HIP events (misohip_profile_*) over a few steps;
- batches (`tune_ensemble`): the y/z sweep walk length per run size, timed
  on a probe batch of that size.

The state is restored after tuning, so tuning changes no result (every
candidate computes the same values; only the launch shape differs).
Choices are cached in a JSON file ($MISO_AMD_TILE_CACHE, default
~/.cache/miso_amd/tiles.json) keyed by device, kernel and size.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np

from . import hip

RK_PENCILS = ((1, 4), (2, 4), (2, 8), (4, 2), (4, 4), (4, 8), (4, 16), (8, 4), (8, 8))
CHUNKS = (16, 32, 64, 128, 256)
ZCHUNKS = (16, 32, 64, 128, 256, 512)


def _device_key() -> str:
    import torch
    p = torch.cuda.get_device_properties(0)
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}|{p.multi_processor_count}"


class TileCache:
    def __init__(self, path: str | None = None):
        self.path = path or os.environ.get("MISO_AMD_TILE_CACHE") or os.path.join(
            os.path.expanduser("~"), ".cache", "miso_amd", "tiles.json")
        self.data = {}
        try:
            with open(self.path) as f:
                self.data = json.load(f)
        except (OSError, ValueError):
            self.data = {}

    def get(self, key):
        return self.data.get(key)

    def put(self, key, value) -> None:
        self.data[key] = value
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump(self.data, f, indent=1, sort_keys=True)
        except OSError:
            pass  # an unwritable cache only costs a re-tune


def get_tiles(ctx):
    t = hip.MisohipTiles()
    hip.call("misohip_get_tiles", ctx.handle, C.byref(t))
    return t


def set_tiles(ctx, pencil_bx: int, pencil_by: int, chunk: int, variant: int = 0,
              zchunk: int = 32, fuse_xy: int = 0, xy_chunk: int = 64) -> None:
    t = hip.MisohipTiles()
    t.rk_pencil_bx, t.rk_pencil_by, t.sweep_chunk = int(pencil_bx), int(pencil_by), int(chunk)
    t.rk_variant, t.rk_zchunk, t.av_fuse_xy = int(variant), int(zchunk), int(fuse_xy)
    t.av_xy_chunk = int(xy_chunk)
    hip.call("misohip_set_tiles", ctx.handle, C.byref(t))


def _profile(ctx, fn, totals=False):
    """Run fn with the library's event timing on; returns {kind: ms per
    launch} (totals: {kind: summed ms})."""
    L = hip.lib()
    nk = len(hip.KIND_NAMES)
    hip.check(L.misohip_profile_enable(ctx.handle, 1))
    fn()
    ms = (C.c_double * nk)()
    cnt = (C.c_int * nk)()
    hip.check(L.misohip_profile_read(ctx.handle, ms, cnt, nk))
    hip.check(L.misohip_profile_enable(ctx.handle, 0))
    return {hip.KIND_NAMES[k]: ms[k] / (1 if totals else cnt[k]) for k in range(nk) if cnt[k]}


def _timed_rounds(cands, time_one, rounds: int = 2):
    """{candidate: mean time over `rounds` passes}, every other pass in
    reverse order: a candidate timed k-th in one pass is timed k-th from
    the end in the next, so a steady drift of the device's clock over the
    passes adds the same to every candidate's mean."""
    tot = {}
    for r in range(rounds):
        for c in (cands if r % 2 == 0 else cands[::-1]):
            tot[c] = tot.get(c, 0.0) + time_one(c)
    return {c: t / rounds for c, t in tot.items()}


def tune_single(ctx, model, bc, steps: int = 2, pencils=RK_PENCILS, chunks=CHUNKS,
                zchunks=ZCHUNKS, cache: TileCache | None = None, force: bool = False):
    """Choose and set the RK pencil, variant and z chunk and the sweep chunk
    for `model` (a mhd.MHD); returns {"rk_pencil": (bx, by), "rk_variant": v,
    "rk_zchunk": z, "sweep_chunk": c, "times_ms": {...}, "seconds": tuning
    time, "cached": bool}."""
    import torch
    cache = cache or TileCache()
    g = model.grid
    # |v7: entries tuned before the overlapped AV walks existed are not reused
    key = f"{_device_key()}|single|{g.i_size}x{g.j_size}x{g.k_size}|{g.dtype.name}|v7"
    hit = None if force else cache.get(key)
    if hit:
        set_tiles(ctx, *hit["rk_pencil"], hit["sweep_chunk"], hit.get("rk_variant", 0),
                  hit.get("rk_zchunk", 32), hit.get("av_fuse_xy", 0), hit.get("av_xy_chunk", 64))
        return dict(hit, cached=True, seconds=0.0)
    t0 = time.perf_counter()
    saved = model.qq.data.clone()
    base = get_tiles(ctx)
    integ = model.integrator

    def run():
        for _ in range(steps):
            integ.step(bc, model.qq)
        model.qq.data.copy_(saved)

    run()  # warm-up (code objects, caches)
    rk_cands = []
    # 6 (the 12-row z-march) is an FP32 kernel; FP64 would run it as 5; on
    # an x-y grid the z-march variants walk y instead (one candidate)
    xy = g.i_size > 1 and g.j_size > 1 and g.k_size == 1
    variants = ((0, 1, 2, 3, 4, 5) + ((6,) if g.dtype == np.float32 else ())
                if min(g.i_size, g.j_size, g.k_size) > 1 else (0, 4) if xy else (0,))
    for var in variants:
        if var in (4, 5, 6) and xy:
            cands = [(1, 1, 32)]
        elif var in (4, 5, 6):  # the z-march kernels have no pencils; their knob is the z chunk
            cands = [(1, 1, z) for z in zchunks]
        elif var in (0, 3):
            cands = [(bx, by, 32) for bx, by in pencils]
        else:  # the register-capped variants are tried on the default pencil only
            cands = [(2, 8, 32)]
        for bx, by, z in cands:
            rk_cands.append((bx, by, var, z))

    def rk_time(c):
        set_tiles(ctx, c[0], c[1], base.sweep_chunk, c[2], c[3])
        return _profile(ctx, run).get("rk_stage", float("inf"))

    rk_times = _timed_rounds(rk_cands, rk_time)
    best_p = min(rk_times, key=rk_times.get)
    bx, by, var, zc = best_p
    # AV: summed time of the sweeps and ghost fills, separate x/y sweeps for
    # every walk length, then (3D) the one-walk x-y sweeps for every walk
    # length at the best z walk
    kinds = ("av_x", "av_y", "av_z", "ghost_fill")
    z_times = {}

    def sep_time(c):
        set_tiles(ctx, bx, by, c[1], var, zc)
        t = _profile(ctx, run, totals=True)
        z_times[c[1]] = min(z_times.get(c[1], float("inf")), t.get("av_z", 0.0))
        return sum(t.get(k, 0.0) for k in kinds)

    sweep_times = _timed_rounds([(0, c, 0) for c in chunks], sep_time)
    if min(g.i_size, g.j_size, g.k_size) > 1:
        cz = min(z_times, key=z_times.get)

        def xy_time(c):
            set_tiles(ctx, bx, by, c[1], var, zc, c[0], c[2])
            t = _profile(ctx, run, totals=True)
            return sum(t.get(k, 0.0) for k in kinds)

        cands = [(fz, cz, xc) for fz in ((1, 2) if g.dtype == np.float64 else (1,))
                 for xc in chunks]
        if bc.periodic[2] in (1, 2):
            # 3 and 4: 1 and 2 with the z walk overlapped with the x-y walk
            # (a periodic or fixed z axis), over z walk lengths that give it
            # a few slabs; the x-y walk lengths around the usual choice
            cands += [(fz, c, xc) for fz in ((3, 4) if g.dtype == np.float64 else (3,))
                      for c in sorted({cz, 16, 32, 64, 128}) for xc in (32, 64, 128) if xc in chunks]
        sweep_times.update(_timed_rounds(cands, xy_time))
    fz, best_c, xc = min(sweep_times, key=sweep_times.get)
    xc = xc or 64
    # the three fastest RK candidates again, over whole steps with the chosen
    # sweeps: the RK launch shape changes what the next pass finds in the
    # caches (FP32 512^3: the same x-y walk took 4.6 ms after one z-march
    # shape and 5.0 ms after another)
    top = sorted(rk_times, key=rk_times.get)[:3]

    def step_time(c):
        set_tiles(ctx, c[0], c[1], best_c, c[2], c[3], fz, xc)
        return sum(_profile(ctx, run, totals=True).values())

    step_times = _timed_rounds(top, step_time)
    bx, by, var, zc = min(step_times, key=step_times.get)
    set_tiles(ctx, bx, by, best_c, var, zc, fz, xc)
    torch.cuda.synchronize()
    out = {"rk_pencil": [bx, by], "rk_variant": int(var), "rk_zchunk": int(zc),
           "sweep_chunk": int(best_c), "av_fuse_xy": int(fz), "av_xy_chunk": int(xc),
           "times_ms": {"rk_stage": {(f"z{z}v{v}" if v in (4, 5, 6) else f"{a}x{b}v{v}"): t
                                     for (a, b, v, z), t in rk_times.items()},
                        "step_top3": {(f"z{z}v{v}" if v in (4, 5, 6) else f"{a}x{b}v{v}"): t
                                      for (a, b, v, z), t in step_times.items()},
                        "av_and_ghost_per_run": {
                            (f"xy{x}z{c}" + ("w2" if f in (2, 4) else "") +
                             ("o" if f >= 3 else "") if f else f"sep{c}"): v
                            for (f, c, x), v in sweep_times.items()}}}
    cache.put(key, out)
    return dict(out, cached=False, seconds=time.perf_counter() - t0)


def tune_ensemble(ctx, ens, make_probe, chunks=CHUNKS, reps: int = 3,
                  cache: TileCache | None = None, force: bool = False):
    """Choose the y/z sweep walk length per run size of `ens` (an
    ensemble.Ensemble) and apply it.  make_probe(size) builds a small
    Ensemble of runs of that size (the probe is timed, the real runs are
    untouched).  Returns {size: chunk}, the timings and the tuning time."""
    import torch
    cache = cache or TileCache()
    t0 = time.perf_counter()
    sizes = sorted({(r.grid.i_size, r.grid.j_size, r.grid.k_size) for r in ens.runs})
    choice, times = {}, {}
    for s in sizes:
        key = f"{_device_key()}|batch|{s[0]}x{s[1]}x{s[2]}|{ens.dtype.name}"
        hit = None if force else cache.get(key)
        if hit:
            choice[s] = hit["chunk"]
            continue
        probe = make_probe(s)
        tt = {}
        for c in chunks:
            probe.set_chunks([c] * probe.n)
            probe.step(1)  # warm-up
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            probe.step(reps)
            e1.record()
            torch.cuda.synchronize()
            tt[c] = e0.elapsed_time(e1) / reps
        probe.close()
        best = min(tt, key=tt.get)
        choice[s] = int(best)
        times[f"{s[0]}x{s[1]}x{s[2]}"] = {str(k): v for k, v in tt.items()}
        cache.put(key, {"chunk": int(best), "times_ms": times[f"{s[0]}x{s[1]}x{s[2]}"]})
    ens.set_chunks([choice[(r.grid.i_size, r.grid.j_size, r.grid.k_size)] for r in ens.runs])
    return {"chunks": {f"{a}x{b}x{c}": v for (a, b, c), v in choice.items()}, "times_ms": times,
            "seconds": time.perf_counter() - t0}


def summary(d) -> dict:
    """JSON-safe copy of a tuning result."""
    return json.loads(json.dumps(d, default=lambda o: list(o) if isinstance(o, tuple) else str(o)))


__all__ = ["TileCache", "tune_single", "tune_ensemble", "set_tiles", "get_tiles", "summary",
           "RK_PENCILS", "CHUNKS", "ZCHUNKS"]
