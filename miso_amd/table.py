"""Device uniform-table interpolators, the host mirror of MISO's
UniformTableInterpolator1D / 2D (miso/include/miso/table_interpolator.hpp:22-166)
over misohip_table1d / misohip_table2d: a table of n (n0 x n1) values
spanning [x_min, x_max] uniformly, linear (bilinear) interpolation, the
nearest table value outside the range.  Tables, inputs and outputs are
device tensors of one dtype (float32 or float64); inputs of any shape
(flat or 3D fields) are interpolated element by element."""
from __future__ import annotations

import ctypes as C

from . import hip


def _dtype_code(t):
    import torch
    if t.dtype == torch.float64:
        return hip.F64
    if t.dtype == torch.float32:
        return hip.F32
    raise TypeError("tables hold float32 or float64 values")


def _check(t, dtype, device, what):
    if t.dtype != dtype or t.device != device or not t.is_contiguous():
        raise ValueError(f"{what} must be a contiguous {dtype} tensor on {device}")


class UniformTableInterpolator1D:
    def __init__(self, ctx, table, x_min: float, x_max: float):
        if table.dim() != 1 or table.numel() < 2 or not x_max > x_min:
            raise ValueError("a 1D table needs at least two entries and x_max > x_min")
        self.ctx, self.table = ctx, table.contiguous()
        self.x_min, self.x_max = float(x_min), float(x_max)
        self.code = _dtype_code(self.table)

    def interpolate(self, x, out=None):
        import torch
        _check(x, self.table.dtype, self.table.device, "x")
        y = torch.empty_like(x) if out is None else out
        _check(y, self.table.dtype, self.table.device, "out")
        if y.numel() != x.numel():
            raise ValueError("x and out differ in size")
        hip.call("misohip_table1d", self.ctx.handle, self.code, C.c_void_p(self.table.data_ptr()),
                 int(self.table.numel()), self.x_min, self.x_max, C.c_void_p(x.data_ptr()),
                 C.c_void_p(y.data_ptr()), int(x.numel()))
        return y

    __call__ = interpolate


class UniformTableInterpolator2D:
    def __init__(self, ctx, table, x0_min: float, x0_max: float, x1_min: float, x1_max: float):
        if table.dim() != 2 or min(table.shape) < 2 or not (x0_max > x0_min and x1_max > x1_min):
            raise ValueError("a 2D table needs at least 2 x 2 entries and increasing ranges")
        self.ctx, self.table = ctx, table.contiguous()
        self.r = (float(x0_min), float(x0_max), float(x1_min), float(x1_max))
        self.code = _dtype_code(self.table)

    def interpolate(self, x0, x1, out=None):
        import torch
        for t, w in ((x0, "x0"), (x1, "x1")):
            _check(t, self.table.dtype, self.table.device, w)
        y = torch.empty_like(x0) if out is None else out
        _check(y, self.table.dtype, self.table.device, "out")
        if not (x0.numel() == x1.numel() == y.numel()):
            raise ValueError("x0, x1 and out differ in size")
        n0, n1 = self.table.shape
        hip.call("misohip_table2d", self.ctx.handle, self.code, C.c_void_p(self.table.data_ptr()),
                 int(n0), int(n1), *self.r, C.c_void_p(x0.data_ptr()), C.c_void_p(x1.data_ptr()),
                 C.c_void_p(y.data_ptr()), int(x0.numel()))
        return y

    __call__ = interpolate


__all__ = ["UniformTableInterpolator1D", "UniformTableInterpolator2D"]
