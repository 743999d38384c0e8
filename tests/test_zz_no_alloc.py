"""misohip_update_noalloc: the update with its workspace in the library's
per-device scratch array instead of the caller's misohip_work.

Checked against misohip_update, the entry point it varies, through the C ABI:
the same initial state advanced by one update of each gives the same bits
(fields and ghosts), on two calls in a row, the second on a smaller grid (so
it runs on what the first left in the scratch array), in FP64 and FP32; and a
grid whose workspace does not fit is refused with MISOHIP_E_RANGE before any
work is queued.
"""
import ctypes as C

import numpy as np
import pytest

from miso_amd import hip, problems

E_RANGE = -5


def _cases():
    # (problem, sizes, dtype): the first larger than the second
    return [(problems.turbulence, (40, 36, 32), np.float64),
            (problems.orszag_tang, (48, 40, 1), np.float64),
            (problems.turbulence, (44, 40, 36), np.float32),
            (problems.turbulence, (20, 16, 12), np.float32)]


def _setup(ctx, fn, sizes, dtype):
    from miso_amd import mhd
    g = mhd.Grid(*sizes, 2, dtype=dtype, device="cuda:0")
    cfg = mhd.Config({"mhd": {"cfl_number": 0.4}})
    integ = mhd.Integrator(cfg, g, ctx, gm=5.0 / 3.0)
    bc = mhd.BoundaryCondition((True, True, True))
    q0 = problems.to_numpy(fn(g)).astype(dtype)
    return g, integ, bc, q0


def _fields(ctx, g, q0):
    from miso_amd import mhd
    F = mhd.Fields(g)
    F.copy_from_host(ctx, np.ascontiguousarray(q0))
    return F


@pytest.mark.gpu
def test_update_noalloc_equals_update(ctx):
    for fn, sizes, dtype in _cases():
        print(f"noalloc vs update: {fn.__name__} {sizes} {np.dtype(dtype).name}", flush=True)
        g, integ, bc, q0 = _setup(ctx, fn, sizes, dtype)
        a = _fields(ctx, g, q0)
        b = _fields(ctx, g, q0)
        integ.apply_boundary_condition(bc, a)
        integ.apply_boundary_condition(bc, b)
        dt = integ.cfl(a)
        integ.update(dt, bc, a)
        hip.call("misohip_update_noalloc", ctx.handle, C.byref(g.c_struct()), integ.dtype_code,
                 C.byref(integ.params), C.byref(bc.c_struct()), C.byref(b.c_struct()),
                 float(dt))
        ha, hb = a.to_host(ctx), b.to_host(ctx)
        # compare the bit patterns as a boolean first: a failing == on two
        # multi-megabyte bytes objects makes pytest build a diff of them
        same = bool(np.array_equal(ha.view(np.uint8), hb.view(np.uint8)))
        where = [] if same else np.argwhere(ha.view(np.uint8) != hb.view(np.uint8))[:4].tolist()
        assert same, f"{fn.__name__} {sizes}: noalloc differs at (field, i, j, k) bytes {where}"
        assert not np.array_equal(ha, q0), "the update changed nothing"


@pytest.mark.gpu
def test_update_noalloc_with_the_av_walks_overlapped():
    """With the AV walks overlapped (av_fuse_xy 4, a periodic box) on a
    fresh context: the first no-alloc call takes the sequential walks (it
    creates no streams), misohip_update then the overlapped ones, and a
    second no-alloc call the overlapped ones on the streams the update made;
    all three give misohip_update's bits."""
    from miso_amd import mhd
    from miso_amd.autotune import set_tiles
    c = mhd.ExecContext(0)
    try:
        set_tiles(c, 1, 1, 8, 5, 32, 4, 16)  # 8-plane z chunks: four slabs at 32 planes
        g, integ, bc, q0 = _setup(c, problems.turbulence, (40, 36, 32), np.float64)
        outs = []
        for which in ("noalloc", "update", "noalloc"):
            F = _fields(c, g, q0)
            integ.apply_boundary_condition(bc, F)
            dt = 2.5e-3
            if which == "update":
                integ.update(dt, bc, F)
            else:
                hip.call("misohip_update_noalloc", c.handle, C.byref(g.c_struct()),
                         integ.dtype_code, C.byref(integ.params), C.byref(bc.c_struct()),
                         C.byref(F.c_struct()), float(dt))
            outs.append(F.to_host(c))
        for n, h in enumerate(outs[1:], 1):
            same = bool(np.array_equal(h.view(np.uint8), outs[0].view(np.uint8)))
            assert same, f"call {n} differs from the first"
        assert not np.array_equal(outs[0], q0), "the update changed nothing"
    finally:
        c.close()


@pytest.mark.gpu
def test_update_noalloc_refuses_grids_beyond_the_scratch(ctx):
    from miso_amd import mhd
    g = mhd.Grid(72, 72, 72, 2, dtype=np.float64, device="cuda:0")
    integ = mhd.Integrator(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, gm=5.0 / 3.0)
    bc = mhd.BoundaryCondition((True, True, True))
    q = mhd.Fields(g)
    rc = hip.lib().misohip_update_noalloc(ctx.handle, C.byref(g.c_struct()), integ.dtype_code,
                                          C.byref(integ.params), C.byref(bc.c_struct()),
                                          C.byref(q.c_struct()), 1e-3)
    assert rc == E_RANGE
