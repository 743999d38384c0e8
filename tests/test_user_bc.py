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


@pytest.mark.gpu
def test_python_user_boundary_matches_descriptor(ctx):
    """The Python mirror's user boundary ("user" axis, a function of the
    (9, Kt, Jt, It) tensor being filled, torch operations on the stream):
    a z mirror with odd vz and bz written with torch gives the mirror
    descriptor's bits over a 3-step run, every field and dt."""
    import torch

    from miso_amd import mhd, problems
    g = mhd.Grid(20, 16, 12, 2, device="cuda:0")
    q0 = problems.turbulence(g, device="cpu", seed=5)
    odd = (0, 0, 0, 1, 0, 0, 1, 0, 0)
    sign = torch.tensor([-1.0 if o else 1.0 for o in odd], dtype=torch.float64,
                        device="cuda:0").view(9, 1, 1)
    m_, kt = g.k_margin, g.k_total

    def z_mirror(t):
        for layer in range(m_):
            t[:, layer] = sign * t[:, 2 * m_ - 1 - layer]
            t[:, kt - m_ + layer] = sign * t[:, kt - m_ - 1 - layer]

    signs = ((1,) * 9, (1,) * 9, tuple(-1 if o else 1 for o in odd))
    runs = []
    for bc in (mhd.BoundaryCondition((True, True, False), sym_sign=signs),
               mhd.BoundaryCondition((True, True, "user"), user=z_mirror)):
        m = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, 5.0 / 3.0)
        m.qq.load_ijk(q0)
        m.integrator.apply_boundary_condition(bc, m.qq)
        d = m.integrator.run(bc, m.qq, 3)
        torch.cuda.synchronize()
        runs.append((d, m.qq.data.clone()))
    assert runs[0][0] == runs[1][0]
    assert torch.equal(runs[0][1], runs[1][1])
    with pytest.raises(ValueError):
        mhd.BoundaryCondition((True, True, "user"))
