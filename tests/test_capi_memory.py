"""Raw device memory through the C ABI: misohip_alloc / misohip_memcpy /
misohip_free (the storage and copies of the reference's Array3D on the
device, miso/include/miso/array3d.hpp:156-167,206-211,254-263)."""
import ctypes as C

import numpy as np
import pytest

from miso_amd import hip


def test_argument_errors_without_gpu():
    L = hip.lib()
    p = C.c_void_p()
    assert L.misohip_alloc(None, 8, C.byref(p)) == hip.E_ARG
    assert L.misohip_free(None, None) == hip.E_ARG
    assert L.misohip_memcpy(None, None, None, 0, hip.MEMCPY_D2D) == hip.E_ARG


@pytest.mark.gpu
def test_alloc_copy_round_trip(ctx):
    L = hip.lib()
    n = 1 << 20
    src = np.random.default_rng(3).integers(0, 256, n, dtype=np.uint8)
    back = np.zeros(n, dtype=np.uint8)
    a, b, z = C.c_void_p(), C.c_void_p(), C.c_void_p(1)
    hip.call("misohip_alloc", ctx.handle, n, C.byref(a))
    hip.call("misohip_alloc", ctx.handle, n, C.byref(b))
    hip.call("misohip_alloc", ctx.handle, 0, C.byref(z))
    assert a.value and b.value and z.value is None
    try:
        hip.call("misohip_memcpy", ctx.handle, a, src.ctypes.data, n, hip.MEMCPY_H2D)
        hip.call("misohip_memcpy", ctx.handle, b, a, n, hip.MEMCPY_D2D)
        hip.call("misohip_memcpy", ctx.handle, back.ctypes.data, b, n, hip.MEMCPY_D2H)
        hip.call("misohip_synchronize", ctx.handle)
        assert np.array_equal(back, src)
        # a partial copy at an offset, and the empty copy
        part = np.zeros(1000, dtype=np.uint8)
        hip.call("misohip_memcpy", ctx.handle, part.ctypes.data, C.c_void_p(b.value + 4096),
                 1000, hip.MEMCPY_D2H)
        hip.call("misohip_memcpy", ctx.handle, None, None, 0, hip.MEMCPY_D2H)
        hip.call("misohip_synchronize", ctx.handle)
        assert np.array_equal(part, src[4096:5096])
        assert hip.lib().misohip_memcpy(ctx.handle, back.ctypes.data, b, n, 7) == hip.E_ARG
    finally:
        hip.call("misohip_free", ctx.handle, a)
        hip.call("misohip_free", ctx.handle, b)
        hip.call("misohip_free", ctx.handle, None)
