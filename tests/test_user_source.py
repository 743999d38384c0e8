"""User source terms (Source::vx, vy, vz, ei; mhd_model_base.hpp:131-155).

The reference inlines the functor into the RK stage, evaluated on argm at
every interior cell (mhd_integrator_host.hpp:134,151,168,241).  Here a
source callback writes the four terms of each stage's argm into arrays that
the stage kernel adds at the same places (misohip_set_source_callback,
mhd.UserSource).  Checks: constant gravity written by a user function gives
the built-in gravity descriptor's bits; the point-mass gravity of
demo/mhd3d_magnetosphere/src/src.hpp:18-34, with a heating src.ei = h ro,
gives the oracle's bits (fields and dts) over several steps.
"""
import numpy as np
import pytest

from miso_amd import problems
from tests.parity_util import assert_bitwise

pytestmark = pytest.mark.gpu


def _coords(g):
    import torch
    dev = "cuda:0"
    x = torch.as_tensor(np.asarray(g.x), dtype=torch.float64, device=dev).view(1, 1, -1)
    y = torch.as_tensor(np.asarray(g.y), dtype=torch.float64, device=dev).view(1, -1, 1)
    z = torch.as_tensor(np.asarray(g.z), dtype=torch.float64, device=dev).view(-1, 1, 1)
    return x, y, z


def test_user_gravity_equals_the_gravity_descriptor(ctx):
    import torch

    from miso_amd import mhd
    from miso_amd.rmhd import Stratification, stratified_box
    grav = Stratification().grav
    sizes = (16, 16, 64)
    bounds = ((0.0, 0.25), (0.0, 0.25), (0.0, 1.0))

    def user_gravity(argm, out):
        out[0].zero_()
        out[1].zero_()
        torch.mul(argm[0], -grav, out=out[2])
        out[3].zero_()

    res = []
    for src in (dict(src_kind=1, grav_axis=2, grav=grav),
                dict(source=mhd.UserSource(user_gravity))):
        g = mhd.Grid(*sizes, 2, bounds=bounds, device="cuda:0")
        q0 = problems.to_numpy(stratified_box(g, Stratification(), seed=4))
        fx = mhd.Fields(g)
        fx.copy_from_host(ctx, q0)
        bc = mhd.BoundaryCondition((1, 1, 2), fixed=fx)
        m = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, 5.0 / 3.0, **src)
        m.apply_initial_condition(q0, bc)
        dts = m.integrator.run(bc, m.qq, 3) + [m.integrator.step(bc, m.qq) for _ in range(2)]
        torch.cuda.synchronize()
        res.append((dts, m.qq.data.clone()))
    assert res[0][0] == res[1][0]
    assert torch.equal(res[0][1], res[1][1])


@pytest.mark.parametrize("use_run", [False, True])
def test_point_mass_source_matches_the_oracle(ctx, use_run):
    import torch

    from miso_amd import mhd
    from oracle import oracle as O
    sizes, nsteps = (20, 16, 12), 4
    bounds = ((1.0, 2.0), (-0.5, 0.5), (0.5, 1.5))  # the point mass sits outside the box
    g_grav, heat = 0.3, 0.05
    og = O.make_grid(sizes, 2, bounds=bounds)
    g = mhd.Grid(*sizes, 2, bounds=bounds, device="cuda:0")
    q0 = problems.to_numpy(problems.turbulence(g, seed=12))
    op = O.Params(cfl_number=0.4, src_kind=2, grav=g_grav, src_heat=heat)
    ref = q0.copy()
    O.ghost_fill(og, op, ref)
    ref_dts = O.run(og, op, ref, nsteps)

    x, y, z = _coords(g)
    r3 = None

    def point_mass(argm, out):
        nonlocal r3
        if r3 is None:  # pow3(sqrt(x^2 + y^2 + z^2)), as the demo writes it
            r = torch.sqrt(x * x + y * y + z * z)
            r3 = (r * r) * r
        ro = argm[0]
        for a, xa in enumerate((x, y, z)):
            out[a] = (-ro) * g_grav * xa / r3
        out[3] = heat * ro

    m = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, 5.0 / 3.0,
                source=mhd.UserSource(point_mass))
    bc = mhd.BoundaryCondition((1, 1, 1))
    m.apply_initial_condition(q0, bc)
    if use_run:
        dts = m.integrator.run(bc, m.qq, nsteps)
    else:
        dts = [m.integrator.step(bc, m.qq) for _ in range(nsteps)]
    out = m.qq.to_host(ctx)
    sl = (slice(None),) + g.interior()
    assert np.array_equal(np.array(dts), ref_dts), (dts, ref_dts)
    assert_bitwise(out[sl], ref[sl], f"point-mass source, run={use_run}")
    # the source mattered: without it the oracle ends elsewhere
    plain = q0.copy()
    op0 = O.Params(cfl_number=0.4)
    O.ghost_fill(og, op0, plain)
    O.run(og, op0, plain, nsteps)
    assert not np.array_equal(plain[sl], ref[sl])


@pytest.mark.parametrize("use_run", [0, 1])
def test_cpp_make_source_matches_the_oracle(ctx, use_run):
    """The same point-mass Source written as a MISO Source struct (device
    members vx, vy, vz, ei) and run through miso::hip::make_source and the
    C++ MHD class (tests/hip/test_user_bc.hip): the oracle's bits."""
    import ctypes as C

    from miso_amd import mhd
    from oracle import oracle as O
    from tests.test_user_bc import _lib
    D = C.POINTER(C.c_double)
    L = _lib()
    f = L.hip_user_source_point_mass
    f.argtypes = [C.c_int, C.c_int, C.c_int, D, D, C.c_int, C.c_int, C.c_double, C.c_double, D,
                  D]
    f.restype = C.c_int
    sizes, nsteps = (20, 16, 12), 4
    bounds = ((1.0, 2.0), (-0.5, 0.5), (0.5, 1.5))
    g_grav, heat = 0.3, 0.05
    og = O.make_grid(sizes, 2, bounds=bounds)
    g = mhd.Grid(*sizes, 2, bounds=bounds, device="cpu")
    q0 = np.ascontiguousarray(problems.to_numpy(problems.turbulence(g, seed=12)))
    ref = q0.copy()
    op = O.Params(cfl_number=0.4, src_kind=2, grav=g_grav, src_heat=heat)
    O.ghost_fill(og, op, ref)
    ref_dts = O.run(og, op, ref, nsteps)
    b = np.array([v for ab in bounds for v in ab], dtype=np.float64)
    out = np.zeros_like(q0)
    dts = np.zeros(nsteps)
    rc = f(*sizes, b.ctypes.data_as(D), q0.ctypes.data_as(D), nsteps, use_run, g_grav, heat,
           out.ctypes.data_as(D), dts.ctypes.data_as(D))
    assert rc == 0
    sl = (slice(None),) + g.interior()
    assert np.array_equal(dts, ref_dts), (dts, ref_dts)
    assert_bitwise(out[sl], ref[sl], f"make_source point mass, run={use_run}")
