"""Per-column MLP surrogate of the radiative heating Q_rad (no counterpart in
the reference).

One column is the vertical line of nz interior cells at (i, j).  Its inputs
are (ln ro, ln T, |B|^2) over the column (3 nz values, normalised per
feature and height) and, with features=4, the horizontal Laplacian of ln T
(what a column sees of its neighbours' temperature, which inclined rays
carry in); its output is Q_rad over the column (nz values, normalised per
height).  The network (features x nz -> hidden x layers -> nz, SiLU)
lives in libmisohip.so: BF16 weights on the MFMA units with FP32
accumulation, FP32 master weights, Adam, all forward and backward kernels
hand-written (misohip_mlp_*).  torch only holds device buffers and shuffles
sample indices.

Training data comes from the GPU RT solver: `collect_columns` runs the
radiative box with a plain RT solve on every step and keeps, at chosen
steps, the raw column features and the solved Q_rad columns.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import hip


def interior_columns(field, grid):
    """(Kt, Jt, It) device field -> (ncol, nz) with col = j * ni + i."""
    k0, j0, i0 = grid.k_margin, grid.j_margin, grid.i_margin
    sub = field[k0:k0 + grid.k_size, j0:j0 + grid.j_size, i0:i0 + grid.i_size]
    return sub.permute(1, 2, 0).reshape(grid.i_size * grid.j_size, grid.k_size)


class ColumnSurrogate:
    def __init__(self, ctx, nz: int, hidden: int = 256, hidden_layers: int = 3,
                 max_batch: int = 4096, seed: int = 0, device="cuda", features: int = 3):
        import torch
        assert features in (3, 4)
        self.ctx, self.nz, self.device = ctx, nz, device
        self.n_features = features
        self.n_in, self.n_out = features * nz, nz
        self.sizes = [self.n_in] + [hidden] * hidden_layers + [self.n_out]
        self.max_batch = int(max_batch)
        h = C.c_void_p()
        hip.call("misohip_mlp_create", ctx.handle, self.n_in, hidden, hidden_layers, self.n_out,
                 self.max_batch, C.byref(h))
        self.h = h
        rng = np.random.default_rng(seed)
        for l, (i, o) in enumerate(zip(self.sizes[:-1], self.sizes[1:])):
            w = (rng.standard_normal((o, i)) * math.sqrt(2.0 / i)).astype(np.float32)
            if l == len(self.sizes) - 2:
                w *= 0.1
            b = np.zeros(o, dtype=np.float32)
            hip.call("misohip_mlp_set_layer", ctx.handle, h, l, w.ctypes.data_as(C.c_void_p),
                     b.ctypes.data_as(C.c_void_p))
        self.set_normalization(np.zeros(self.n_in), np.ones(self.n_in), np.zeros(self.n_out),
                               np.ones(self.n_out))
        self._X = None
        self._Y = None
        self.history = []

    def close(self):
        if self.h:
            hip.lib().misohip_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- normalisation --------------------------------------------------------
    def set_normalization(self, mu_in, sd_in, mu_out, sd_out):
        self.norm = [np.ascontiguousarray(a, dtype=np.float32) for a in (mu_in, sd_in, mu_out, sd_out)]
        p = [a.ctypes.data_as(C.POINTER(C.c_float)) for a in self.norm]
        hip.call("misohip_mlp_set_norm", self.ctx.handle, self.h, *p)

    def _buffers(self, ncol):
        import torch
        if self._X is None or self._X.shape[0] < ncol:
            self._X = torch.empty((ncol, self.n_in), dtype=torch.bfloat16, device=self.device)
            self._Xf = torch.empty((ncol, self.n_in), dtype=torch.float32, device=self.device)
            self._Y = torch.empty((ncol, self.n_out), dtype=torch.float32, device=self.device)
        return self._X, self._Xf, self._Y

    # -- features and inference ----------------------------------------------
    def features(self, grid, qq, gm, keep_float=False):
        """Normalised column features of the fields into the bf16 buffer
        (and an fp32 copy when keep_float)."""
        ncol = grid.i_size * grid.j_size
        X, Xf, _ = self._buffers(ncol)
        hip.call("misohip_mlp_features", self.ctx.handle, self.h, C.byref(grid.c_struct()),
                 C.byref(qq.c_struct()), float(gm), C.c_void_p(X.data_ptr()),
                 C.c_void_p(Xf.data_ptr()) if keep_float else None)
        return X[:ncol], Xf[:ncol]

    def predict(self, grid, qq, gm, qrad_out, flags=None, q_rt=None, corr=None):
        """Q_rad of every interior cell into qrad_out (a (Kt, Jt, It) float64
        field), plus the correction field corr when given; columns with
        flags != 0 take q_rt."""
        ncol = grid.i_size * grid.j_size
        X, _ = self.features(grid, qq, gm)
        _, _, Y = self._buffers(ncol)
        hip.call("misohip_mlp_qrad", self.ctx.handle, self.h, C.byref(grid.c_struct()),
                 C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()),
                 C.c_void_p(flags.data_ptr()) if flags is not None else None,
                 C.c_void_p(q_rt.data_ptr()) if q_rt is not None else None,
                 C.c_void_p(corr.data_ptr()) if corr is not None else None,
                 C.c_void_p(qrad_out.data_ptr()))
        return qrad_out

    def residual(self, grid, q_rt, flags, tol, corr=None):
        """Flag the columns whose last prediction misses q_rt by more than tol
        (relative L2 over the column); corr (optional field) receives
        q_rt - prediction."""
        _, _, Y = self._buffers(grid.i_size * grid.j_size)
        hip.call("misohip_mlp_residual", self.ctx.handle, self.h, C.byref(grid.c_struct()),
                 C.c_void_p(Y.data_ptr()), C.c_void_p(q_rt.data_ptr()),
                 C.c_void_p(flags.data_ptr()), float(tol),
                 C.c_void_p(corr.data_ptr()) if corr is not None else None)

    # -- training -------------------------------------------------------------
    def fit(self, feats, targets, epochs=40, batch=1024, lr=1e-3, lr_min=1e-5, val_frac=0.1,
            seed=0, log_every=0, log=None):
        """Fit on raw features (N, 3 nz) fp32 and raw Q_rad columns (N, nz)
        (device tensors): sets the normalisation from the training split,
        then runs Adam with a cosine learning-rate schedule.  Returns
        (train MSE of the last epoch, validation relative L2 of Q_rad)."""
        import torch
        n = feats.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        perm = torch.randperm(n, generator=g).to(feats.device)
        nval = max(batch, int(n * val_frac)) if n > 2 * batch else 0
        tr, va = perm[nval:], perm[:nval]
        ntr = (len(tr) // batch) * batch
        f64 = feats[tr].double()
        t64 = targets[tr].double()
        mu_in = f64.mean(0)
        sd_in = torch.maximum(f64.std(0), 1e-6 * f64.abs().mean(0)).clamp_min(1e-12)
        mu_out = t64.mean(0)
        sd_out = t64.std(0).clamp_min(1e-12)
        self.set_normalization(mu_in.cpu().numpy(), sd_in.cpu().numpy(), mu_out.cpu().numpy(),
                               sd_out.cpu().numpy())
        Xn = ((feats.double() - mu_in) / sd_in).to(torch.bfloat16)
        Yn = ((targets.double() - mu_out) / sd_out).float()
        del f64, t64
        steps = epochs * (ntr // batch)
        k = 0
        last = 0.0
        for ep in range(epochs):
            order = tr[torch.randperm(len(tr), generator=g).to(feats.device)][:ntr]
            Xe, Ye = Xn[order].contiguous(), Yn[order].contiguous()
            tot = 0.0
            for b0 in range(0, ntr, batch):
                loss = C.c_double()
                want = (b0 + batch >= ntr)
                hip.call("misohip_mlp_grad", self.ctx.handle, self.h,
                         C.c_void_p(Xe[b0:b0 + batch].data_ptr()),
                         C.c_void_p(Ye[b0:b0 + batch].data_ptr()), batch,
                         C.byref(loss) if want else None)
                lr_k = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * k / max(1, steps)))
                hip.call("misohip_mlp_adam", self.ctx.handle, self.h, C.c_float(lr_k),
                         C.c_float(0.9), C.c_float(0.999), C.c_float(1e-8), C.c_float(0.0))
                k += 1
                if want:
                    tot = loss.value
            last = tot
            self.history.append(tot)
            if log_every and log and (ep % log_every == 0 or ep == epochs - 1):
                log(f"  epoch {ep:3d} train mse {tot:.3e}")
        val = None
        if nval:
            val = self.validate(Xn[va], targets[va], mu_out, sd_out)
        return last, val

    def forward(self, Xn):
        """Normalised network output of normalised bf16 inputs (N, 3 nz)."""
        import torch
        N = Xn.shape[0]
        Y = torch.empty((N, self.n_out), dtype=torch.float32, device=Xn.device)
        for b0 in range(0, N, self.max_batch):
            nb = min(self.max_batch, N - b0)
            hip.call("misohip_mlp_forward", self.ctx.handle, self.h,
                     C.c_void_p(Xn[b0:b0 + nb].contiguous().data_ptr()), nb,
                     C.c_void_p(Y[b0:b0 + nb].data_ptr()))
        return Y

    def validate(self, Xn, targets, mu_out, sd_out) -> float:
        Y = self.forward(Xn).double() * sd_out + mu_out
        t = targets.double()
        return float((Y - t).norm() / t.norm())


def collect_columns(box, seeds, steps, sample_every, max_columns=None, rng_seed=0, log=None,
                    features=3):
    """Run the radiative box with a plain RT solve on every step for each
    seed and keep (raw features, Q_rad) columns at steps 0, sample_every, ...
    below `steps`.  With max_columns, an even random subsample of that many
    columns over all snapshots is kept.  Returns device tensors
    (N, features nz) fp32 and (N, nz) float64."""
    import torch
    grid = box.grid
    ncol = grid.i_size * grid.j_size
    snaps = [s for s in range(steps) if s % sample_every == 0]
    nsnap = len(seeds) * len(snaps)
    per = ncol if max_columns is None else min(ncol, -(-max_columns // nsnap))
    gen = torch.Generator(device="cpu").manual_seed(rng_seed)
    feats, targs = [], []
    raw = _RawFeatures(box.ctx, grid, features)
    for seed in seeds:
        box.initialize(seed)
        box.steps_done = 0
        for s in range(steps):
            q = box.solve_rt()
            if s % sample_every == 0:
                sel = torch.randperm(ncol, generator=gen)[:per].to(q.device)
                Xf = raw(box.model.qq, box.strat.gm)
                feats.append(Xf[sel].clone())
                targs.append(interior_columns(q, grid)[sel].clone())
                if log:
                    log(f"  seed {seed} step {s}: rt iterations {box.rt_iters[-1]}")
            box._advance(q)
    raw.close()
    F = torch.cat(feats)
    T = torch.cat(targs)
    if max_columns is not None and F.shape[0] > max_columns:
        F, T = F[:max_columns], T[:max_columns]
    return F, T


class _RawFeatures:
    """Un-normalised column features through misohip_mlp_features (a small
    network handle with identity normalisation)."""

    def __init__(self, ctx, grid, features=3):
        import torch
        self.ctx, self.grid = ctx, grid
        nz = grid.k_size
        nin = features * nz
        h = C.c_void_p()
        hip.call("misohip_mlp_create", ctx.handle, nin, 32, 1, nz, 32, C.byref(h))
        self.h = h
        z = np.zeros(nin, np.float32)
        o = np.ones(nin, np.float32)
        p = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        self._keep = (z, o)
        hip.call("misohip_mlp_set_norm", ctx.handle, h, p(z), p(o), p(z[:nz]), p(o[:nz]))
        ncol = grid.i_size * grid.j_size
        self.X = torch.empty((ncol, nin), dtype=torch.bfloat16, device=grid.device)
        self.Xf = torch.empty((ncol, nin), dtype=torch.float32, device=grid.device)

    def __call__(self, qq, gm):
        hip.call("misohip_mlp_features", self.ctx.handle, self.h, C.byref(self.grid.c_struct()),
                 C.byref(qq.c_struct()), float(gm), C.c_void_p(self.X.data_ptr()),
                 C.c_void_p(self.Xf.data_ptr()))
        return self.Xf

    def close(self):
        if self.h:
            hip.lib().misohip_mlp_destroy(self.h)
            self.h = None
