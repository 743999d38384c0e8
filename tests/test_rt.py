"""Radiative transfer: pins of the RT oracle (CPU) and parity of the GPU
solver with it (GPU).

MISO has no RT value tests; its demo is the 2D searchlight
(demo/rt2d_searchlight).  Pins: the reference's own quadrature header
(oracle/_ref), closed forms (uniform thermal medium gives I = S and
Q_rad = 0; a single ray through a source-free absorber decays as
exp(-kappa s) exactly in the short-characteristics integral).
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O


def _bounds():
    return ((-0.5, 0.5),) * 3


def test_quadrature_matches_reference_header():
    from tests.test_oracle_pins import lib_ref
    L = lib_ref()
    L.ref_quadrature_f64.argtypes = [C.c_int, C.POINTER(C.c_double)]
    L.ref_quadrature_f64.restype = C.c_int
    out = np.zeros(96)
    n = L.ref_quadrature_f64(24, out.ctypes.data_as(C.POINTER(C.c_double)))
    ref = out[:4 * n].reshape(n, 4)
    w, mx, my, mz = O.rt_quadrature(24)
    assert np.array_equal(ref[:, 0], w) and np.array_equal(ref[:, 1], mx)
    assert np.array_equal(ref[:, 2], my) and np.array_equal(ref[:, 3], mz)
    # Carlson A4: unit directions, weights sum to one, every octant once per group
    assert np.allclose(mx ** 2 + my ** 2 + mz ** 2, 1.0, rtol=0, atol=1e-15)
    assert abs(w.sum() - 1.0) < 1e-15


def test_uniform_thermal_medium():
    """kappa, S uniform with thermal (I = S) inflow on every face: I = S on
    every swept node and Q_rad = 0."""
    g = O.make_grid((12, 10, 8), 2, bounds=_bounds())
    kap = np.full(g.shape, 2.0)
    S = np.full(g.shape, 0.7)
    I = np.zeros((24,) + g.shape)
    it, md = O.rt_solve(g, kap, S, I, periodic=(0, 0, 0), face_kind=(1,) * 6, tolerance=1e-12,
                        max_iters=20)
    q = O.rt_qrad(g, kap, S, I)
    lat = (slice(1, -2), slice(1, -2), slice(1, -2))
    assert np.allclose(I[(slice(None),) + lat], 0.7, rtol=1e-14)
    assert np.abs(q[lat]).max() < 1e-12


def test_single_ray_absorption_is_exponential():
    """One +z ray (the non-A4 quadrature), S = 0, uniform kappa, I = 1 on
    the inner z face: I(z_n) = exp(-kappa n dz) exactly per segment."""
    g = O.make_grid((1, 1, 32), 2, bounds=_bounds())
    kap = np.full(g.shape, 3.0)
    S = np.zeros(g.shape)
    I = np.zeros((1,) + g.shape)
    O.rt_solve(g, kap, S, I, periodic=(0, 0, 0), face_kind=(0, 0, 0, 0, 3, 0),
               face_value=(0, 0, 0, 0, 1.0, 0), tolerance=1e-14, max_iters=5, num_rays=1)
    dz = g.axis("dz")[0]
    n = np.arange(0, 33)
    expect = np.exp(-3.0 * dz * n)
    assert np.allclose(I[0, 0, 0, 1:34], expect, rtol=1e-13)


# ------------------------------------------------------------------ GPU parity
def _gpu_rt(ctx, og, kap, S, num_rays, **kw):
    import torch

    from miso_amd import mhd
    from miso_amd.rt import RT
    sizes = (og.g.i_size, og.g.j_size, og.g.k_size)
    g = mhd.Grid(*sizes, 2, bounds=_bounds(), device="cuda:0")
    rt = RT(g, num_rays)
    rt.abs_coeff.copy_(RT.to_dev_field(kap, "cuda:0"))
    rt.src_func.copy_(RT.to_dev_field(S, "cuda:0"))
    it, md = rt.solve(ctx, **kw)
    q = rt.qrad(ctx, **kw)
    torch.cuda.synchronize()
    I = rt.intensities().permute(0, 3, 2, 1).cpu().numpy()
    return it, md, I, RT.to_host_field(q)


@pytest.mark.gpu
def test_gpu_searchlight_2d(ctx):
    """demo/rt2d_searchlight set-up: 64x64, kappa = 1, S = 0.5, a beam of
    radius 0.2 entering through the inner x face, y and z periodic."""
    og = O.make_grid((64, 64, 1), 2, bounds=_bounds())
    kap = np.ones(og.shape)
    S = np.full(og.shape, 0.5)
    # beam: incoming intensity 1 inside r <= 0.2 on the x-inner face; the
    # constant-face boundary is the searchlight's top hat through S at the face
    I_ref = np.zeros((24,) + og.shape)
    kw = dict(periodic=(0, 1, 1), face_kind=(3, 0, 0, 0, 0, 0), face_value=(1.0, 0, 0, 0, 0, 0),
              tolerance=1e-5, max_iters=100)
    it_ref, md_ref = O.rt_solve(og, kap, S, I_ref, **kw)
    q_ref = O.rt_qrad(og, kap, S, I_ref)
    it, md, I, q = _gpu_rt(ctx, og, kap, S, 24, periodic=(False, True, True), face_kind=kw["face_kind"],
                           face_value=kw["face_value"], tolerance=1e-5, max_iters=100)
    assert it == it_ref
    assert abs(md - md_ref) <= 1e-6 * md_ref + 1e-12
    assert np.allclose(I, I_ref, rtol=1e-13, atol=1e-15)
    assert np.allclose(q, q_ref, rtol=1e-12, atol=1e-13)


@pytest.mark.gpu
def test_gpu_stratified_3d(ctx):
    """Stratified box: x, y periodic, thermal inflow at the bottom, vacuum at
    the top, kappa and S varying in all three directions."""
    og = O.make_grid((16, 12, 20), 2, bounds=_bounds())
    rng = np.random.default_rng(7)
    z = og.axis("z")[None, None, :]
    T = 2.0 - 1.5 * (z + 0.5) + 0.02 * rng.standard_normal(og.shape)
    ro = np.exp(-(z + 0.5) / 0.1) * (1 + 0.02 * rng.standard_normal(og.shape))
    kap = 50.0 * ro * T ** 3
    S = T ** 4
    I_ref = np.zeros((24,) + og.shape)
    kw = dict(periodic=(1, 1, 0), face_kind=(0, 0, 0, 0, 1, 2), tolerance=1e-9, max_iters=300)
    it_ref, md_ref = O.rt_solve(og, kap, S, I_ref, **kw)
    q_ref = O.rt_qrad(og, kap, S, I_ref)
    it, md, I, q = _gpu_rt(ctx, og, kap, S, 24, periodic=(True, True, False),
                           face_kind=kw["face_kind"], tolerance=1e-9, max_iters=300)
    assert it == it_ref, (it, it_ref)
    scale = np.abs(I_ref).max()
    assert np.abs(I - I_ref).max() <= 1e-12 * scale
    lat = (slice(1, -2), slice(1, -2), slice(1, -2))
    assert np.abs(q[lat] - q_ref[lat]).max() <= 1e-11 * np.abs(q_ref[lat]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("sizes", [(64, 48, 6), (40, 72, 9)])
def test_gpu_thin_slab_per_ray_sweep(ctx, sizes):
    """A thin slab (x, y periodic, thermal bottom, vacuum top; the learned
    RT's bottom-slab shape): the z rays end after a few planes, the x- and
    y-dominant rays sweep on alone; against the oracle."""
    og = O.make_grid(sizes, 2, bounds=_bounds())
    rng = np.random.default_rng(11)
    z = og.axis("z")[None, None, :]
    T = 1.5 - 0.5 * (z + 0.5) + 0.05 * rng.standard_normal(og.shape)
    kap = 20.0 * (1 + 0.1 * rng.standard_normal(og.shape)) ** 2
    S = T ** 4
    I_ref = np.zeros((24,) + og.shape)
    kw = dict(periodic=(1, 1, 0), face_kind=(0, 0, 0, 0, 1, 2), tolerance=1e-9, max_iters=300)
    it_ref, md_ref = O.rt_solve(og, kap, S, I_ref, **kw)
    q_ref = O.rt_qrad(og, kap, S, I_ref)
    it, md, I, q = _gpu_rt(ctx, og, kap, S, 24, periodic=(True, True, False),
                           face_kind=kw["face_kind"], tolerance=1e-9, max_iters=300)
    assert it == it_ref, (it, it_ref)
    scale = np.abs(I_ref).max()
    assert np.abs(I - I_ref).max() <= 1e-12 * scale
    lat = (slice(1, -2), slice(1, -2), slice(1, -2))
    assert np.abs(q[lat] - q_ref[lat]).max() <= 1e-11 * np.abs(q_ref[lat]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("sizes,periodic,faces", [
    ((24, 20, 16), (True, True, False), (0, 0, 0, 0, 1, 2)),
    ((20, 28, 12), (False, False, False), (1, 3, 2, 1, 1, 2)),
    ((36, 30, 1), (False, True, True), (3, 1, 0, 0, 0, 0)),
    ((72, 40, 6), (True, True, False), (0, 0, 0, 0, 1, 2)),
    ((40, 70, 31), (False, True, False), (2, 0, 0, 0, 3, 1)),
])
def test_gpu_ray_layouts_and_multi_plane_sweeps_same_bits(ctx, sizes, periodic, faces):
    """Keeping the x-dominant rays y-fastest, and sweeping the rays of a
    short c axis several planes per launch, change where intensities live
    and how the sweep is launched, not what it computes: iterations, last
    change, every intensity and Q_rad equal bit for bit to the all
    x-fastest, one-plane-per-launch solve, over periodic and face
    boundaries of every kind on x and y."""
    import torch

    from miso_amd import mhd
    from miso_amd.rt import RT
    g = mhd.Grid(*sizes, 2, bounds=((-0.5, 0.5), (-0.4, 0.6), (-0.5, 0.5)), device="cuda:0")
    rng = np.random.default_rng(5)
    shp = (g.k_total, g.j_total, g.i_total)
    kap = torch.from_numpy(5.0 * (1 + 0.2 * rng.random(shp))).to("cuda:0")
    S = torch.from_numpy(1 + 0.3 * rng.random(shp)).to("cuda:0")
    kw = dict(periodic=periodic, face_kind=faces, face_value=(0.8, 1.3, 0, 0, 0, 0),
              tolerance=1e-10, max_iters=200)
    out = []
    for yf, planes in ((False, 0), (True, 0), (True, 8), (False, 5), (True, 16), (True, 2)):
        rt = RT(g, 24)
        rt.abs_coeff.copy_(kap)
        rt.src_func.copy_(S)
        it, md = rt.solve(ctx, x_rays_yfast=yf, sweep_planes=planes, **kw)
        q = rt.qrad(ctx, x_rays_yfast=yf, sweep_planes=planes, **kw).clone()
        torch.cuda.synchronize()
        out.append((it, md, rt.yfast_mask, rt.intensities().cpu(), q.cpu()))
    it0, md0, m0, I0, q0 = out[0]
    assert m0 == 0
    # the rays kept y-fastest are those whose shortest step to the next
    # plane is along x (one cell on z: no step along z)
    w, mx, my, mz = rt.quadrature()
    d = [1.0 / sizes[0], 1.0 / sizes[1], 1.0 / sizes[2] if sizes[2] > 1 else 1e30]
    dom = [int(np.argmin([d[a] / max(abs(m[n]), 1e-30) for a, m in enumerate((mx, my, mz))]))
           for n in range(24)]
    assert out[1][2] == sum(1 << n for n in range(24) if dom[n] == 0)
    assert out[1][2] & 0xFF == 0xFF
    for it1, md1, _, I1, q1 in out[1:]:
        assert (it0, md0) == (it1, md1)
        assert torch.equal(I0, I1)
        assert torch.equal(q0, q1)


@pytest.mark.gpu
def test_gpu_single_ray_and_option_checks(ctx):
    """The one-ray quadrature (a +z ray) through an absorber with a constant
    incoming face, on a grid whose y axis is short enough for the
    multi-plane sweep: against the oracle (intensities to 1e-13), the same
    bits with one plane per launch, and sweep_planes outside 0..16 refused."""
    import torch

    from miso_amd import hip, mhd
    from miso_amd.rt import RT
    og = O.make_grid((10, 6, 40), 2, bounds=_bounds())
    rng = np.random.default_rng(2)
    kap = 2.0 + rng.random(og.shape)
    S = 0.3 * rng.random(og.shape)
    kw = dict(periodic=(1, 1, 0), face_kind=(0, 0, 0, 0, 3, 2), face_value=(0, 0, 0, 0, 1.0, 0),
              tolerance=1e-12, max_iters=50)
    I_ref = np.zeros((1,) + og.shape)
    it_ref, _ = O.rt_solve(og, kap, S, I_ref, num_rays=1, **kw)
    q_ref = O.rt_qrad(og, kap, S, I_ref, num_rays=1)
    g = mhd.Grid(10, 6, 40, 2, bounds=_bounds(), device="cuda:0")
    outs = []
    for planes in (16, 0):
        rt = RT(g, 1)
        rt.abs_coeff.copy_(RT.to_dev_field(kap, "cuda:0"))
        rt.src_func.copy_(RT.to_dev_field(S, "cuda:0"))
        it, md = rt.solve(ctx, sweep_planes=planes, **kw)
        q = rt.qrad(ctx, sweep_planes=planes, **kw)
        torch.cuda.synchronize()
        outs.append((it, rt.intensities().permute(0, 3, 2, 1).cpu().numpy(), RT.to_host_field(q)))
    (it1, I1, q1), (it0, I0, q0) = outs
    assert it1 == it0 == it_ref
    assert np.array_equal(I1, I0) and np.array_equal(q1, q0)
    assert np.allclose(I1, I_ref, rtol=1e-13, atol=1e-15)
    lat = (slice(1, -2), slice(1, -2), slice(1, -2))
    assert np.allclose(q1[lat], q_ref[lat], rtol=1e-12, atol=1e-13)
    with pytest.raises(hip.MisohipError):
        RT(g, 1).solve(ctx, sweep_planes=17, **kw)
