"""User boundary functors (Boundary::USER in include/miso_hip.hpp, the
device helpers of include/miso_hip_device.hpp, misohip_set_bc_callback):
MISO's BC::apply runs before the halo exchange
(mhd_integrator_host.hpp:250-255) and a device boundary is a functor that
calls for_each / bc::symmetric (boundary_condition.hpp:108-182,
demo/mhd3d_magnetosphere/src/bc.hpp:39-158).  The same run with a boundary
descriptor and with a user functor writing the same ghost values must give
the same bits: every field, ghosts included, and every dt."""
import ctypes as C
import os

import numpy as np
import pytest

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hip", "_build",
                   "libmiso_hip_user_bc.so")
D = C.POINTER(C.c_double)


def _lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB)
    for name in ("hip_user_bc_mirror_z", "hip_user_bc_fixed_x"):
        f = getattr(L, name)
        f.argtypes = [C.c_int, C.c_int, C.c_int, D, C.c_int, D, D, D, D]
        f.restype = C.c_int
    return L


def test_user_bc_library_exports():
    L = _lib()
    assert L.hip_user_bc_mirror_z and L.hip_user_bc_fixed_x


def _state(sizes, seed):
    """A 3D turbulence-like state in MISO order with ghosts (9, It, Jt, Kt)."""
    from miso_amd import mhd, problems
    g = mhd.Grid(*sizes, 2, bounds=((0.0, 1.0),) * 3, dtype=np.float64, device="cpu")
    q = problems.turbulence(g, device="cpu", seed=seed)
    return np.ascontiguousarray(q.cpu().numpy() if hasattr(q, "cpu") else q, dtype=np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["hip_user_bc_mirror_z", "hip_user_bc_fixed_x"])
def test_user_functor_matches_descriptor(ctx, case):
    sizes, nsteps = (20, 16, 12), 3
    q = _state(sizes, 5)
    a, b = np.zeros_like(q), np.zeros_like(q)
    da, db = np.zeros(nsteps), np.zeros(nsteps)
    rc = getattr(_lib(), case)(*sizes, q.ctypes.data_as(D), nsteps, a.ctypes.data_as(D),
                               b.ctypes.data_as(D), da.ctypes.data_as(D), db.ctypes.data_as(D))
    assert rc == 0
    assert np.array_equal(da, db) and (da > 0).all()
    assert np.array_equal(a, b)
    assert not np.array_equal(a, q)  # the run moved the state
