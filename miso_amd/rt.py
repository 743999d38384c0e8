"""Python host mirror of MISO's rt::RT (miso/include/miso/rt_host.hpp:104-689)
over the GPU solver in libmisohip.so.

RT(grid, num_rays) owns the device arrays of the reference's struct --
rint and rint_old (num_rays fields), src_func and abs_coeff -- laid out
x-fastest like the MHD fields, except that with x_rays_yfast (the default)
a ray whose dominant axis is x keeps its field of rint y-fastest, so its
sweep planes are contiguous; `yfast_mask` names those rays as of the last
solve, and intensities() gives every ray as (k, j, i).
solve() is RT::solve with the boundary set given per face (periodic axes
are self-exchanged); qrad() is the radiative heating rate
Q_rad = 4 pi kappa (J - S) this backend adds.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip


class RT:
    def __init__(self, grid, num_rays: int = 24):
        import torch
        assert grid.dtype == np.float64, "the RT solver runs in FP64"
        self.grid = grid
        self.num_rays = 24 if num_rays == 24 else 1
        shp = (grid.k_total, grid.j_total, grid.i_total)
        dev = grid.device
        self.rint = torch.zeros((self.num_rays,) + shp, dtype=torch.float64, device=dev)
        self.rint_old = torch.zeros_like(self.rint)
        self.src_func = torch.zeros(shp, dtype=torch.float64, device=dev)
        self.abs_coeff = torch.zeros(shp, dtype=torch.float64, device=dev)
        self.qrad_field = torch.zeros(shp, dtype=torch.float64, device=dev)
        self.yfast_mask = 0
        self._layouts = {}                   # x_rays_yfast -> mask (the grid is fixed)

    def config(self, periodic=(True, True, False), face_kind=(0, 0, 0, 0, 1, 2),
               face_value=(0.0,) * 6, tolerance=1e-5, max_iters=100,
               x_rays_yfast=True, sweep_planes=16) -> hip.MisohipRtConfig:
        c = hip.MisohipRtConfig()
        c.num_rays = self.num_rays
        for a in range(3):
            c.periodic[a] = 1 if periodic[a] else 0
        for f in range(6):
            c.face_kind[f] = int(face_kind[f])
            c.face_value[f] = float(face_value[f])
        c.tolerance = float(tolerance)
        c.max_iters = int(max_iters)
        c.x_rays_yfast = 1 if x_rays_yfast else 0
        c.sweep_planes = int(sweep_planes)
        return c

    def ray_layout(self, ctx, **kw) -> int:
        """Bit mask of the rays whose field of rint is y-fastest under the
        configuration `kw` (as given to solve)."""
        m = C.c_uint()
        hip.call("misohip_rt_ray_layout", ctx.handle, C.byref(self.grid.c_struct()),
                 C.byref(self.config(**kw)), C.byref(m))
        return int(m.value)

    def intensities(self):
        """rint with every ray's field as (k, j, i), x fastest (a copy when
        the last solve kept rays y-fastest)."""
        mask = self.yfast_mask
        if not mask:
            return self.rint
        out = self.rint.clone()
        kt, jt, it = self.rint.shape[1:]
        for n in range(self.num_rays):
            if (mask >> n) & 1:
                out[n] = self.rint[n].reshape(kt, it, jt).permute(0, 2, 1)
        return out

    def solve(self, ctx, **kw):
        """RT::solve; returns (iterations, last max change)."""
        cfg = self.config(**kw)
        yf = bool(kw.get("x_rays_yfast", True))
        if yf not in self._layouts:
            self._layouts[yf] = self.ray_layout(ctx, **kw)
        self.yfast_mask = self._layouts[yf]
        it, md = C.c_int(), C.c_double()
        hip.call("misohip_rt_solve", ctx.handle, C.byref(self.grid.c_struct()), C.byref(cfg),
                 C.c_void_p(self.abs_coeff.data_ptr()), C.c_void_p(self.src_func.data_ptr()),
                 C.c_void_p(self.rint.data_ptr()), C.c_void_p(self.rint_old.data_ptr()),
                 C.byref(it), C.byref(md))
        return it.value, md.value

    def quadrature(self):
        """(weights, mu_x, mu_y, mu_z) of the solver's rays
        (AngularQuadrature, rt_quadrature.hpp:14-63)."""
        bufs = [(C.c_double * 24)() for _ in range(4)]
        n = hip.lib().misohip_rt_quadrature(self.num_rays, *bufs)
        if n < 0:
            hip.check(n)
        return tuple(np.array(b[:n]) for b in bufs)

    def qrad(self, ctx, **kw):
        cfg = self.config(**kw)
        hip.call("misohip_rt_qrad", ctx.handle, C.byref(self.grid.c_struct()), C.byref(cfg),
                 C.c_void_p(self.abs_coeff.data_ptr()), C.c_void_p(self.src_func.data_ptr()),
                 C.c_void_p(self.rint.data_ptr()), C.c_void_p(self.qrad_field.data_ptr()))
        return self.qrad_field

    def source_from_fields(self, ctx, qq, gm: float, kappa0: float, s_rad: float = 1.0) -> None:
        """kappa = kappa0 ro T^3, S = s_rad T^4 with T = (gm-1) ei, every cell."""
        hip.call("misohip_rt_source", ctx.handle, C.byref(self.grid.c_struct()),
                 C.byref(qq.c_struct()), float(gm), float(kappa0), float(s_rad),
                 C.c_void_p(self.abs_coeff.data_ptr()), C.c_void_p(self.src_func.data_ptr()))

    def heat(self, ctx, qq, dt: float, qrad=None) -> None:
        """Operator-split radiative heating ei += dt*Q_rad/ro (interior), from
        qrad_field or the given field of the same shape."""
        q = self.qrad_field if qrad is None else qrad
        assert q.shape == self.qrad_field.shape and q.dtype == self.qrad_field.dtype
        hip.call("misohip_rt_heat", ctx.handle, C.byref(self.grid.c_struct()),
                 C.byref(qq.c_struct()), C.c_void_p(q.data_ptr()), float(dt))

    # host <-> device helpers in MISO order
    @staticmethod
    def to_dev_field(a: np.ndarray, device):
        import torch
        return torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).to(device)

    @staticmethod
    def to_host_field(t) -> np.ndarray:
        return np.ascontiguousarray(t.permute(2, 1, 0).cpu().numpy())
