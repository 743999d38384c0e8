"""Parity of the HIP backend with the CPU oracle, function by function and
step by step, through the C ABI (include/misohip.h).

Bar: FP64 results equal the oracle's bit for bit on the interior (signed
zeros compare equal); the multi-step runs are also held to the stated
1e-12 relative L-infinity.  Ghost cells are compared on faces only: MISO's
halo exchange leaves edges and corners stale and no stencil reads them
(miso/include/miso/mhd_halo_exchange.hpp:58-180).
"""
import functools

import numpy as np
import pytest

from miso_amd import problems
from miso_amd.rmhd import Stratification, stratified_box
from tests.parity_util import assert_bitwise, random_state, rel_linf

pytestmark = pytest.mark.gpu

GEOMS = {
    "2d": (40, 36, 1),
    "3d": (20, 18, 16),
    "1dx": (48, 1, 1),
    "1dy": (1, 48, 1),
    "1dz": (1, 1, 48),
    "2d_xz": (24, 1, 20),
}


def _grids(sizes, dtype=np.float64):
    from miso_amd import mhd
    from oracle import oracle as O
    og = O.make_grid(sizes, 2, dtype=dtype)
    g = mhd.Grid(*sizes, 2, dtype=dtype, device="cuda:0")
    for n in ("x", "y", "z", "dx", "dy", "dz", "dxi", "dyi", "dzi"):
        assert np.array_equal(og.axis(n), getattr(g, n)), n
    assert og.min_dxyz == g.min_dxyz
    return og, g


def _face_mask(g):
    It, Jt, Kt = g.shape
    gi = np.zeros(It, bool); gi[:g.i_margin] = gi[It - g.i_margin:] = True
    gj = np.zeros(Jt, bool); gj[:g.j_margin] = gj[Jt - g.j_margin:] = True
    gk = np.zeros(Kt, bool); gk[:g.k_margin] = gk[Kt - g.k_margin:] = True
    n = gi[:, None, None].astype(int) + gj[None, :, None] + gk[None, None, :]
    return n <= 1


def _bcs(kind):
    from miso_amd import mhd
    from oracle import oracle as O
    if kind == "periodic":
        return mhd.BoundaryCondition((True, True, True)), dict(periodic=(1, 1, 1))
    if kind == "fixed_z":
        # the caller provides the fixed-value fields (see test_ghost_fill)
        return None, dict(periodic=(1, 1, 2))
    # mirror on y and z with antisymmetric normal velocity / field
    signs = [[1] * 9 for _ in range(3)]
    signs[1][2] = signs[1][5] = -1
    signs[2][3] = signs[2][6] = -1
    return (mhd.BoundaryCondition((True, False, False), tuple(tuple(s) for s in signs)),
            dict(periodic=(1, 0, 0), sym_sign=tuple(tuple(s) for s in signs)))


def _to_dev(ctx, g, q):
    from miso_amd import mhd
    F = mhd.Fields(g)
    F.copy_from_host(ctx, np.ascontiguousarray(q))
    return F


def _integrator(ctx, g, **kw):
    from miso_amd import mhd
    cfg = mhd.Config({"mhd": {"cfl_number": kw.pop("cfl", 0.4),
                              "artificial_viscosity": kw.pop("av", {})}})
    return mhd.Integrator(cfg, g, ctx, gm=kw.pop("gm", 5.0 / 3.0), **kw)


@pytest.mark.parametrize("geom", list(GEOMS))
@pytest.mark.parametrize("kind", ["periodic", "mirror", "fixed_z"])
def test_ghost_fill(ctx, geom, kind):
    from oracle import oracle as O
    sizes = GEOMS[geom]
    og, g = _grids(sizes)
    bc, pk = _bcs(kind)
    q = random_state(g.shape, seed=11)
    if kind == "fixed_z":
        from miso_amd import mhd
        fx = random_state(g.shape, seed=12)
        pk = dict(pk, fixed=fx)
        fixed_dev = _to_dev(ctx, g, fx)
        bc = mhd.BoundaryCondition((True, True, "fixed"), fixed=fixed_dev)
    ref = q.copy()
    O.ghost_fill(og, O.Params(**pk), ref)
    F = _to_dev(ctx, g, q)
    it = _integrator(ctx, g)
    it.apply_boundary_condition(bc, F)
    out = F.to_host(ctx)
    mask = _face_mask(g)
    assert_bitwise(out[:, mask], ref[:, mask], f"ghost fill {geom}/{kind}")
    if kind == "fixed_z" and g.k_size > 1:
        # the fixed axis' ghost planes hold the fixed values, and their edges
        # with the x and y ghosts wrap (only x-y corners are stale in the
        # reference's pack-all-then-unpack exchange)
        It, Jt, Kt = g.shape
        gi = np.zeros(It, bool); gi[:g.i_margin] = gi[It - g.i_margin:] = True
        gj = np.zeros(Jt, bool); gj[:g.j_margin] = gj[Jt - g.j_margin:] = True
        keep = ~(gi[:, None, None] & gj[None, :, None]) & np.ones(g.shape, bool)
        assert_bitwise(out[:, keep], ref[:, keep], f"ghost fill {geom}/{kind} with z edges")


def test_ghost_fill_reference_kat(ctx):
    """The reference's own index checks (miso/tests/serial/
    test_boundary_condition_common.hpp:15-120): on a 10x11x12 grid with
    margin 2, a mirror with sign +1 makes ghost 0 equal cell 3 and ghost 1
    equal cell 2 (and It-1 ~ It-4, It-2 ~ It-3); sign -1 negates."""
    from miso_amd import mhd
    g = mhd.Grid(10, 11, 12, 2, bounds=((0, 1), (0, 2), (0, 3)), device="cuda:0")
    It = g.i_total
    q = np.zeros((9,) + g.shape)
    q[0] = np.arange(It, dtype=np.float64)[:, None, None]
    for sign in (1, -1):
        signs = [[1] * 9 for _ in range(3)]
        signs[0][0] = sign
        bc = mhd.BoundaryCondition((False, True, True), tuple(tuple(s) for s in signs))
        F = _to_dev(ctx, g, q)
        _integrator(ctx, g).apply_boundary_condition(bc, F)
        ro = F.to_host(ctx)[0]
        jk = (slice(2, -2), slice(2, -2))
        assert np.array_equal(ro[(0,) + jk], sign * ro[(3,) + jk])
        assert np.array_equal(ro[(1,) + jk], sign * ro[(2,) + jk])
        assert np.array_equal(ro[(It - 1,) + jk], sign * ro[(It - 4,) + jk])
        assert np.array_equal(ro[(It - 2,) + jk], sign * ro[(It - 3,) + jk])


@pytest.mark.parametrize("geom", list(GEOMS))
@pytest.mark.parametrize("grav", [False, True])
def test_rk_stage(ctx, geom, grav):
    from oracle import oracle as O
    sizes = GEOMS[geom]
    og, g = _grids(sizes)
    src = dict(src_kind=1, grav_axis=2 if sizes[2] > 1 else 1, grav=0.7) if grav else {}
    op = O.Params(**src)
    orgn = random_state(g.shape, seed=1)
    argm = random_state(g.shape, seed=2)
    O.ghost_fill(og, op, orgn)
    O.ghost_fill(og, op, argm)
    dt_s, ch2, tau = 1.3e-3, 2.7, 8.0e-3
    ref = np.zeros_like(orgn)
    O.rk_stage(og, op, dt_s, ch2, tau, orgn, argm, ref)
    it = _integrator(ctx, g, **src)
    Fo, Fa, Fr = _to_dev(ctx, g, orgn), _to_dev(ctx, g, argm), _to_dev(ctx, g, np.zeros_like(orgn))
    it.update_sc4(dt_s, tau, ch2, Fo, Fa, Fr)
    out = Fr.to_host(ctx)
    sl = (slice(None),) + g.interior()
    assert_bitwise(out[sl], ref[sl], f"rk stage {geom}")


@pytest.mark.parametrize("geom", ["2d", "3d", "1dy"])
def test_char_vel(ctx, geom):
    import ctypes as C

    import torch

    from miso_amd import hip
    from oracle import oracle as O
    og, g = _grids(GEOMS[geom])
    q = random_state(g.shape, seed=3)
    ref = O.char_vel(og, O.Params(cs_fac=0.8, ca_fac=1.1, vv_fac=0.9), q)
    it = _integrator(ctx, g, av={"cs_fac": 0.8, "ca_fac": 1.1, "vv_fac": 0.9})
    F = _to_dev(ctx, g, q)
    cc = torch.zeros(F.data.shape[1:], dtype=torch.float64, device="cuda:0")
    hip.call("misohip_char_vel", ctx.handle, C.byref(g.c_struct()), hip.F64,
             C.byref(it.params), C.byref(F.c_struct()), C.c_void_p(cc.data_ptr()))
    torch.cuda.synchronize()
    out = cc.permute(2, 1, 0).cpu().numpy()
    assert_bitwise(out, ref, "char_vel")


@pytest.mark.parametrize("geom", list(GEOMS))
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_av_sweep(ctx, geom, direction):
    import torch

    from oracle import oracle as O
    og, g = _grids(GEOMS[geom])
    av = {"ep": 1.5, "fh": 0.8, "cs_fac": 0.8, "ca_fac": 0.8, "vv_fac": 0.8}
    op = O.Params(**av)
    q = random_state(g.shape, seed=4 + direction, amp=0.5)
    O.ghost_fill(og, op, q)
    cc = O.char_vel(og, op, q)
    dt = 2.1e-3
    ref = np.zeros_like(q)
    O.av_sweep(og, op, direction, dt, q, ref, cc)
    it = _integrator(ctx, g, av=av)
    F, Fr = _to_dev(ctx, g, q), _to_dev(ctx, g, np.zeros_like(q))
    cc_dev = torch.from_numpy(np.ascontiguousarray(cc.transpose(2, 1, 0))).to("cuda:0")
    it.av_sweep(direction, dt, F, Fr, cc_dev)
    out = Fr.to_host(ctx)
    sl = (slice(None),) + g.interior()
    assert_bitwise(out[sl], ref[sl], f"av sweep {geom} dir {direction}")
    if direction == 0:
        stored = cc_dev.permute(2, 1, 0).cpu().numpy()
        assert_bitwise(stored[g.interior()], cc[g.interior()], "cc stored by the x sweep")


@pytest.mark.parametrize("geom", list(GEOMS))
def test_cfl(ctx, geom):
    from oracle import oracle as O
    og, g = _grids(GEOMS[geom])
    q = random_state(g.shape, seed=9)
    ref = O.cfl(og, O.Params(cfl_number=0.4), q)
    it = _integrator(ctx, g, cfl=0.4)
    out = it.cfl(_to_dev(ctx, g, q))
    assert out == ref, (out, ref)


def _run_both(ctx, problem, sizes, nsteps, dtype=np.float64, periodic=(1, 1, 1), sym=None,
              cfl=0.4, av=None, gm=5.0 / 3.0, bounds=((0.0, 1.0),) * 3, src=None,
              use_run=False, ghosts=False):
    """Oracle and GPU runs of the same initial state; a fixed axis (2) takes
    its ghost values from the initial state.  use_run: the GPU side advances
    through one misohip_run call (CFL of later steps fused into the last
    sweep) instead of nsteps misohip_step calls.  ghosts: return the whole
    arrays, ghost cells included, instead of the interiors."""
    from miso_amd import mhd, problems
    from oracle import oracle as O
    src = src or {}
    og = O.make_grid(sizes, 2, bounds=bounds, dtype=np.float64)
    g = mhd.Grid(*sizes, 2, bounds=bounds, dtype=dtype, device="cuda:0")
    q0 = problems.to_numpy(problem(g), np.float64)
    fixed = 2 in tuple(periodic)
    op = O.Params(cfl_number=cfl, gm=gm, periodic=periodic, **(av or {}), **src,
                  **({"sym_sign": sym} if sym else {}),
                  **({"fixed": np.ascontiguousarray(q0)} if fixed else {}))
    ref = q0.copy()
    O.ghost_fill(og, op, ref)
    ref_dts = O.run(og, op, ref, nsteps)
    cfg = mhd.Config({"mhd": {"cfl_number": cfl, "artificial_viscosity": av or {}}})
    m = mhd.MHD(cfg, g, ctx, gm=gm, **src)
    fx = None
    if fixed:
        fx = mhd.Fields(g)
        fx.copy_from_host(ctx, np.ascontiguousarray(q0.astype(dtype)))
    bc = mhd.BoundaryCondition(tuple(int(p) for p in periodic), sym, fixed=fx)
    m.apply_initial_condition(q0.astype(dtype), bc)
    if use_run:
        dts = m.integrator.run(bc, m.qq, nsteps)
    else:
        dts = [m.integrator.step(bc, m.qq) for _ in range(nsteps)]
    out = m.qq.to_host(ctx).astype(np.float64)
    sl = (slice(None),) + (() if ghosts else g.interior())
    return out[sl], ref[sl], np.array(dts), ref_dts


STEP_CASES = {
    "orszag_tang_64": (problems.orszag_tang, (64, 64, 1), 20, {}),
    "turbulence_24": (problems.turbulence, (24, 24, 24), 5, {}),
    # 30 planes: the z-march's chunks on a small grid (4 planes) end in a
    # short one, and the walks' chunks do not divide the axes
    "turbulence_20x16x30": (problems.turbulence, (20, 16, 30), 3, {}),
    "kh_48x32_mirror_y": (
        functools.partial(problems.kelvin_helmholtz, seed=3), (48, 32, 1), 10,
        dict(periodic=(1, 0, 1), gm=1.4, bounds=((-0.5, 0.5), (-0.5, 0.5), (-0.5, 0.5)),
             sym=((1,) * 9, (1, 1, -1, 1, 1, -1, 1, 1, 1), (1,) * 9))),
    "stratified_box_fixed_z_16x16x64": (
        lambda g: stratified_box(g, Stratification(), seed=4), (16, 16, 64), 5,
        dict(periodic=(1, 1, 2), bounds=((0.0, 0.25), (0.0, 0.25), (0.0, 1.0)),
             src=dict(src_kind=1, grav_axis=2, grav=Stratification().grav))),
    "stratified_gravity_16": (
        functools.partial(problems.turbulence, v_rms=0.05, b_rms=0.05), (16, 16, 16), 5,
        dict(periodic=(1, 1, 0), src=dict(src_kind=1, grav_axis=2, grav=0.5),
             sym=((1,) * 9, (1,) * 9, (1, 1, 1, -1, 1, 1, -1, 1, 1)))),
}


@pytest.mark.parametrize("case", list(STEP_CASES))
def test_step_fp64(ctx, case):
    problem, sizes, n, kw = STEP_CASES[case]
    out, ref, dts, ref_dts = _run_both(ctx, problem, sizes, n, **kw)
    assert np.array_equal(dts, ref_dts), (dts, ref_dts)
    err = rel_linf(out, ref)
    assert max(err.values()) <= 1e-12, err
    assert_bitwise(out, ref, case)


@pytest.mark.parametrize("case", list(STEP_CASES))
def test_run_fp64(ctx, case):
    """misohip_run (the next step's CFL reduced inside the last AV sweep)
    gives the oracle's fields and dts bit for bit."""
    problem, sizes, n, kw = STEP_CASES[case]
    out, ref, dts, ref_dts = _run_both(ctx, problem, sizes, n, use_run=True, **kw)
    assert np.array_equal(dts, ref_dts), (dts, ref_dts)
    assert_bitwise(out, ref, f"run {case}")


@pytest.mark.parametrize("case", ["turbulence_24", "turbulence_20x16x30",
                                  "stratified_box_fixed_z_16x16x64", "stratified_gravity_16"])
@pytest.mark.parametrize("variant,zchunk", [(4, 5), (4, 32), (5, 5), (5, 32)])
def test_step_fp64_z_march(ctx, case, variant, zchunk):
    """The z-march RK kernels (variants 4 and 5) against the oracle, bit for bit."""
    from miso_amd.autotune import get_tiles, set_tiles
    problem, sizes, n, kw = STEP_CASES[case]
    t0 = get_tiles(ctx)
    default = (t0.rk_pencil_bx, t0.rk_pencil_by, t0.sweep_chunk, t0.rk_variant, t0.rk_zchunk,
               t0.av_fuse_xy, t0.av_xy_chunk)
    set_tiles(ctx, 1, 1, 64, variant, zchunk)
    try:
        out, ref, dts, ref_dts = _run_both(ctx, problem, sizes, n, use_run=True, **kw)
    finally:
        set_tiles(ctx, *default)
    assert np.array_equal(dts, ref_dts), (dts, ref_dts)
    assert_bitwise(out, ref, f"z-march {case} variant {variant} zchunk {zchunk}")


@pytest.mark.parametrize("case", ["turbulence_24", "turbulence_20x16x30",
                                  "stratified_box_fixed_z_16x16x64", "stratified_gravity_16"])
@pytest.mark.parametrize("fuse", [1, 2, 3, 4])
def test_step_fp64_fused_xy_sweeps(ctx, case, fuse):
    """The one-walk x and y AV sweeps (three or two waves per SIMD; 3 and 4:
    with the z walk overlapped with them in slabs where z is periodic, the
    4-plane z chunks of these grids one or two to a slab) against the
    oracle, bit for bit."""
    from miso_amd.autotune import get_tiles, set_tiles
    problem, sizes, n, kw = STEP_CASES[case]
    t0 = get_tiles(ctx)
    default = (t0.rk_pencil_bx, t0.rk_pencil_by, t0.sweep_chunk, t0.rk_variant, t0.rk_zchunk,
               t0.av_fuse_xy, t0.av_xy_chunk)
    set_tiles(ctx, 1, 1, 16, 5, 32, fuse, 16)
    try:
        out, ref, dts, ref_dts = _run_both(ctx, problem, sizes, n, use_run=True, **kw)
    finally:
        set_tiles(ctx, *default)
    assert np.array_equal(dts, ref_dts), (dts, ref_dts)
    assert_bitwise(out, ref, f"fused x-y sweeps {case}")


@pytest.mark.parametrize("sizes,chunk", [((24, 24, 24), 16), ((20, 16, 30), 16),
                                         ((48, 44, 40), 64), ((64, 64, 64), 16)])
def test_overlapped_walks_run_as_a_pipeline(ctx, sizes, chunk):
    """The grids of the overlapped-walk parity cases above and in
    test_parity_headline.py do take the slab pipeline (it records one launch
    for both walks and none for the z walk alone), so those cases check its
    bits, with a periodic and with a fixed z axis; the same grid with a
    mirrored z axis runs the walks in sequence."""
    from miso_amd import mhd
    from miso_amd.autotune import _profile, get_tiles, set_tiles
    t0 = get_tiles(ctx)
    default = (t0.rk_pencil_bx, t0.rk_pencil_by, t0.sweep_chunk, t0.rk_variant, t0.rk_zchunk,
               t0.av_fuse_xy, t0.av_xy_chunk)
    g = mhd.Grid(*sizes, 2, device="cuda:0")
    m = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, 5.0 / 3.0)
    m.qq.load_ijk(problems.turbulence(g, device="cpu"))
    fixed = mhd.Fields(g)
    fixed.data.copy_(m.qq.data)
    set_tiles(ctx, 1, 1, chunk, 5, 32, 4, 16)
    try:
        kinds = {}
        for name, bc in (("periodic", mhd.BoundaryCondition((1, 1, 1))),
                         ("fixed z", mhd.BoundaryCondition((1, 1, 2), fixed=fixed)),
                         ("mirror z", mhd.BoundaryCondition((1, 1, 0)))):
            m.integrator.apply_boundary_condition(bc, m.qq)
            kinds[name] = _profile(ctx, lambda: m.integrator.step(bc, m.qq), totals=True)
    finally:
        set_tiles(ctx, *default)
    for name in ("periodic", "fixed z"):
        assert "av_x" in kinds[name] and "av_z" not in kinds[name], (name, kinds)
    assert "av_z" in kinds["mirror z"], kinds


@pytest.mark.parametrize("case", ["turbulence_24", "stratified_box_fixed_z_16x16x64",
                                  "stratified_gravity_16"])
@pytest.mark.parametrize("fuse", [1, 3])
def test_run_fills_every_ghost_after_its_last_step(ctx, case, fuse):
    """With periodic ghosts read at their sources (z-march RK, one-walk x-y
    sweeps) misohip_run fills the periodic faces after its last step only:
    the whole arrays, ghost cells included, equal those of the same number
    of misohip_step calls, bit for bit."""
    from miso_amd.autotune import get_tiles, set_tiles
    problem, sizes, n, kw = STEP_CASES[case]
    t0 = get_tiles(ctx)
    default = (t0.rk_pencil_bx, t0.rk_pencil_by, t0.sweep_chunk, t0.rk_variant, t0.rk_zchunk,
               t0.av_fuse_xy, t0.av_xy_chunk)
    set_tiles(ctx, 1, 1, 16, 5, 32, fuse, 16)
    try:
        run, _, run_dts, _ = _run_both(ctx, problem, sizes, n, use_run=True, ghosts=True, **kw)
        steps, _, step_dts, _ = _run_both(ctx, problem, sizes, n, ghosts=True, **kw)
    finally:
        set_tiles(ctx, *default)
    assert np.array_equal(run_dts, step_dts), (run_dts, step_dts)
    assert_bitwise(run, steps, f"run vs steps, ghosts included, {case}")


@pytest.mark.parametrize("sizes", [(96, 1, 1), (1, 80, 1), (1, 1, 72), (40, 1, 24)])
def test_run_lower_dimensions(ctx, sizes):
    """misohip_run on 1D and x-z runs, where the CFL stays a separate pass."""
    out, ref, dts, ref_dts = _run_both(ctx, problems.turbulence, sizes, 6, use_run=True)
    assert np.array_equal(dts, ref_dts), (dts, ref_dts)
    assert_bitwise(out, ref, f"run {sizes}")


def test_run_fp32_equals_steps(ctx):
    """FP32: one misohip_run call and the same number of misohip_step calls
    give identical bits and dts."""
    import torch
    from miso_amd import mhd
    for sizes in ((64, 64, 1), (24, 24, 24)):
        res = []
        for use_run in (False, True):
            g = mhd.Grid(*sizes, 2, dtype=np.float32, device="cuda:0")
            m = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, 5.0 / 3.0)
            bc = mhd.BoundaryCondition((1, 1, 1))
            m.apply_initial_condition(problems.to_numpy(problems.orszag_tang(g)).astype(np.float32), bc)
            dts = (m.integrator.run(bc, m.qq, 7) if use_run
                   else [m.integrator.step(bc, m.qq) for _ in range(7)])
            res.append((m.qq.data.clone(), dts))
        assert res[0][1] == res[1][1], (sizes, res[0][1], res[1][1])
        assert torch.equal(res[0][0], res[1][0]), sizes


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_step_shock_tube(ctx, axis):
    """Sod along x, y and z (the reference's end-to-end test geometry) with
    mirror boundaries on every face."""
    from miso_amd import problems
    sizes = [1, 1, 1]
    sizes[axis] = 128
    out, ref, dts, ref_dts = _run_both(
        ctx, lambda g: problems.shock_tube(g, axis=axis), tuple(sizes), 30, periodic=(0, 0, 0),
        gm=1.4, cfl=0.5)
    assert np.array_equal(dts, ref_dts)
    assert_bitwise(out, ref, f"shock tube axis {axis}")


@pytest.mark.parametrize("variant,zchunk,fuse", [(5, 16, 1), (6, 7, 1), (6, 7, 3)])
def test_run_fp32_turbulence_3d(ctx, variant, zchunk, fuse):
    """FP32 3D decaying turbulence (the benchmark's initial-condition family)
    through misohip_run on the FP32 z-march columns the tuner picks (8 rows
    with each cell's products kept beside it, 12 rows), against the FP64
    oracle: every field within 2e-5 relative L-infinity after 30 steps (the
    benchmark's 64^3 x 50 steps measure 3.2e-6), every dt within 1e-5."""
    from miso_amd.autotune import get_tiles, set_tiles
    t0 = get_tiles(ctx)
    default = (t0.rk_pencil_bx, t0.rk_pencil_by, t0.sweep_chunk, t0.rk_variant, t0.rk_zchunk,
               t0.av_fuse_xy, t0.av_xy_chunk)
    set_tiles(ctx, 1, 1, 64, variant, zchunk, fuse, 64)
    try:
        out, ref, dts, ref_dts = _run_both(ctx, problems.turbulence, (48, 44, 40), 30,
                                           dtype=np.float32, use_run=True)
    finally:
        set_tiles(ctx, *default)
    err = rel_linf(out, ref)
    print(f"FP32 3D variant {variant}: max rel linf {max(err.values()):.2e}", flush=True)
    assert max(err.values()) < 2e-5, err
    assert np.allclose(dts, ref_dts, rtol=1e-5), (dts, ref_dts)


def test_step_fp32_orszag_tang(ctx):
    """FP32 storage with FP64 energy recovery against the FP64 oracle."""
    from miso_amd import problems
    out, ref, dts, ref_dts = _run_both(ctx, problems.orszag_tang, (64, 64, 1), 20,
                                       dtype=np.float32)
    err = rel_linf(out, ref)
    # stated tolerance of the FP32 variant after 20 steps of OT 64^2
    assert max(err.values()) < 2e-4, err
    assert np.allclose(dts, ref_dts, rtol=1e-5), (dts, ref_dts)
