"""Batched ensemble of independent 2D/3D runs (no counterpart in the
reference, which runs one model per process, mhd_model_base.hpp:89-117).

`Ensemble` owns one `mhd.MHD` per run (its grid, fields and workspace) and a
`misohip_batch` handle over all of them: one launch per phase advances every
run, each with its own dt computed and kept on the device; `capture(k)`
records k whole steps in a hipGraph that `step()` replays.  Every run
computes exactly what `Integrator.step` computes for it alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip
from .mhd import MHD, BoundaryCondition, Config, Grid

KH_SIZES = (128, 256, 512, 1024)
# Kelvin-Helmholtz runs: x periodic, y mirror with vy and by odd, z degenerate
KH_BC = BoundaryCondition((True, False, True),
                          ((1,) * 9, (1, 1, -1, 1, 1, -1, 1, 1, 1), (1,) * 9))
KH_BOUNDS = ((-0.5, 0.5), (-0.5, 0.5), (-0.5, 0.5))


def kh_sizes(count: int, seed: int):
    """Run sizes n (n x n cells) drawn by seed from {128, 256, 512, 1024}."""
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice(KH_SIZES, size=count)]


class Ensemble:
    def __init__(self, ctx, shapes, initial, bc: BoundaryCondition = KH_BC, dtype=np.float32,
                 cfl_number: float = 0.4, gm: float = 1.4, bounds=KH_BOUNDS, device="cuda:0"):
        """shapes: list of (i_size, j_size, k_size); initial(m, grid) returns the
        float64 (9, It, Jt, Kt) torch tensor of run m in MISO order."""
        self.ctx, self.bc = ctx, bc
        self.dtype = np.dtype(dtype)
        cfg = Config({"mhd": {"cfl_number": cfl_number}})
        self.runs = []
        for m, (ni, nj, nk) in enumerate(shapes):
            g = Grid(ni, nj, nk, 2, bounds=bounds, dtype=dtype, device=device)
            model = MHD(cfg, g, ctx, gm)
            model.qq.load_ijk(initial(m, g))
            model.integrator.apply_boundary_condition(bc, model.qq)
            self.runs.append(model)
        n = len(self.runs)
        self.n = n
        self._grids = (hip.MisohipGrid * n)(*[r.grid.c_struct() for r in self.runs])
        self._qq = (hip.MisohipFields * n)(*[r.qq.c_struct() for r in self.runs])
        self._work = (hip.MisohipWork * n)(*[r.integrator._work for r in self.runs])
        h = C.c_void_p()
        hip.call("misohip_batch_create", ctx.handle, hip.F64 if self.dtype == np.float64 else hip.F32,
                 n, self._grids, self._qq, self._work, C.byref(self.runs[0].integrator.params),
                 C.byref(bc.c_struct()), C.byref(h))
        self.h = h

    @property
    def cells(self) -> int:
        return sum(r.grid.n_interior for r in self.runs)

    def set_chunks(self, chunks) -> None:
        """Walk length of the y/z sweeps per run (drops a captured graph)."""
        arr = (C.c_int * self.n)(*[int(c) for c in chunks])
        hip.call("misohip_batch_set_chunks", self.ctx.handle, self.h, arr)

    def capture(self, steps_per_graph: int) -> None:
        hip.call("misohip_batch_capture", self.ctx.handle, self.h, int(steps_per_graph))

    def step(self, nsteps: int = 1) -> None:
        hip.call("misohip_batch_step", self.ctx.handle, self.h, int(nsteps))

    def state(self):
        """(dt, time, steps, exp_fallback) of every run (synchronises)."""
        dt = (C.c_double * self.n)()
        t = (C.c_double * self.n)()
        s = (C.c_long * self.n)()
        fb = C.c_int()
        hip.call("misohip_batch_state", self.ctx.handle, self.h, dt, t, s, C.byref(fb))
        return np.array(dt), np.array(t), np.array(s), fb.value

    def close(self) -> None:
        if self.h:
            hip.lib().misohip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
