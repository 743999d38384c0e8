"""Tile autotuner: the cache (CPU); on the GPU, tuning restores the state,
picks a candidate, and a tuned launch shape computes the same bits as the
default one."""
import numpy as np
import pytest

from miso_amd.autotune import TileCache


def test_tile_cache_round_trip(tmp_path):
    p = str(tmp_path / "t.json")
    c = TileCache(p)
    assert c.get("k") is None
    c.put("k", {"chunk": 64})
    assert TileCache(p).get("k") == {"chunk": 64}


@pytest.mark.gpu
def test_tune_single_keeps_results(ctx, tmp_path):
    import torch
    from miso_amd import mhd, problems
    from miso_amd.autotune import get_tiles, set_tiles, tune_single
    t0 = get_tiles(ctx)
    default = (t0.rk_pencil_bx, t0.rk_pencil_by, t0.sweep_chunk, t0.rk_variant, t0.rk_zchunk,
               t0.av_fuse_xy, t0.av_xy_chunk)
    g = mhd.Grid(48, 40, 36, 2, device="cuda:0")
    bc = mhd.BoundaryCondition((True, True, True))
    cfg = mhd.Config({"mhd": {"cfl_number": 0.4}})
    m = mhd.MHD(cfg, g, ctx, 5.0 / 3.0)
    m.qq.load_ijk(problems.turbulence(g, device="cpu"))
    m.integrator.apply_boundary_condition(bc, m.qq)
    before = m.qq.data.clone()
    res = tune_single(ctx, m, bc, steps=1, cache=TileCache(str(tmp_path / "c.json")))
    assert torch.equal(m.qq.data, before)
    assert res["sweep_chunk"] in (16, 32, 64, 128, 256) and len(res["rk_pencil"]) == 2
    m.integrator.step(bc, m.qq)
    tuned = m.qq.data.clone()
    m.qq.data.copy_(before)
    set_tiles(ctx, 2, 8, 64)
    m.integrator.step(bc, m.qq)
    assert torch.equal(m.qq.data, tuned)
    # a second call hits the cache
    assert tune_single(ctx, m, bc, cache=TileCache(str(tmp_path / "c.json")))["cached"]
    set_tiles(ctx, *default)


@pytest.mark.gpu
def test_tune_ensemble_keeps_results(ctx, tmp_path):
    import torch
    from miso_amd import problems
    from miso_amd.autotune import tune_ensemble
    from miso_amd.ensemble import Ensemble

    def init(m, g):
        return problems.kelvin_helmholtz(g, seed=4 + m, device="cpu")
    shapes = [(64, 64, 1), (128, 128, 1), (64, 64, 1)]
    a = Ensemble(ctx, shapes, init, dtype=np.float32)
    b = Ensemble(ctx, shapes, init, dtype=np.float32)
    res = tune_ensemble(ctx, b, lambda s: Ensemble(ctx, [s, s], init, dtype=np.float32),
                        cache=TileCache(str(tmp_path / "e.json")))
    assert set(res["chunks"]) == {"64x64x1", "128x128x1"}
    a.step(2)
    b.step(2)
    for m in range(len(shapes)):
        assert torch.equal(a.runs[m].qq.data, b.runs[m].qq.data)
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_rk_variants_and_pencils_compute_the_same_bits(ctx, dtype):
    """Every RK variant, pencil and z chunk gives the same bits, on grids
    with partial x and y tiles (for the 4-, 8- and 12-row columns) and
    uneven z chunks; variant 6 is the FP32 12-row column (FP64 runs it as 5).
    On the x-y grid (z of extent one, large enough for the single-run walk)
    variants 4-6 walk y (60-output strips, 16-row chunks, both cut short)."""
    import numpy as np
    import torch
    from miso_amd import mhd, problems
    from miso_amd.autotune import get_tiles, set_tiles
    t0 = get_tiles(ctx)
    default = (t0.rk_pencil_bx, t0.rk_pencil_by, t0.sweep_chunk, t0.rk_variant, t0.rk_zchunk,
               t0.av_fuse_xy, t0.av_xy_chunk)
    dt = np.float64 if dtype == "f64" else np.float32
    # 1536 x 1313 x 1: 26 strips (the last cut short) x 83 row chunks (the
    # last one row) = 2158 waves, over the single run's y-walk threshold
    for sizes in ((70, 36, 20), (66, 37, 13), (130, 25, 9), (1536, 1313, 1)):
        g = mhd.Grid(*sizes, 2, dtype=dt, device="cuda:0")
        bc = mhd.BoundaryCondition((True, True, True))
        m = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, 5.0 / 3.0)
        m.qq.load_ijk(problems.turbulence(g, device="cpu"))
        m.integrator.apply_boundary_condition(bc, m.qq)
        start = m.qq.data.clone()
        outs = []
        for tiles in ((2, 8, 64, 0), (1, 4, 16, 1), (4, 16, 128, 2), (2, 4, 32, 2), (4, 8, 64, 3),
                      (1, 4, 64, 3), (1, 1, 64, 4, 1), (1, 1, 64, 4, 3), (1, 1, 32, 4, 8),
                      (1, 1, 64, 4, 64), (1, 1, 64, 5, 1), (1, 1, 64, 5, 7), (1, 1, 64, 5, 32),
                      (1, 1, 64, 6, 1), (1, 1, 64, 6, 7), (1, 1, 64, 6, 32)):
            m.qq.data.copy_(start)
            set_tiles(ctx, *tiles)
            m.integrator.step(bc, m.qq)
            outs.append(m.qq.data.clone())
        set_tiles(ctx, *default)
        for n, o in enumerate(outs[1:], 1):
            assert torch.equal(o, outs[0]), (sizes, n)


@pytest.mark.gpu
@pytest.mark.parametrize("ybc", ["periodic", "mirror", "fixed"])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_fused_xy_sweeps_compute_the_same_bits(ctx, ybc, dtype):
    """The x and y AV sweeps as one walk (compiled for three or for two waves
    per SIMD) equal the separate sweeps bit for bit (3D, partial x segments,
    small and uneven y extents and chunks, both z-march RK tiles, which with
    the walk read periodic ghosts at their sources), for periodic, mirrored
    (with sign flips) and fixed y boundaries."""
    import numpy as np
    import torch
    from miso_amd import mhd, problems
    from miso_amd.autotune import get_tiles, set_tiles
    t0 = get_tiles(ctx)
    default = (t0.rk_pencil_bx, t0.rk_pencil_by, t0.sweep_chunk, t0.rk_variant, t0.rk_zchunk,
               t0.av_fuse_xy, t0.av_xy_chunk)
    dt = np.float64 if dtype == "f64" else np.float32
    sym = ((1,) * 9, (1, 1, -1, 1, 1, -1, 1, 1, 1), (1,) * 9)
    for sizes in ((70, 2, 9), (130, 13, 6), (61, 37, 5)):
        g = mhd.Grid(*sizes, 2, dtype=dt, device="cuda:0")
        m = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, 5.0 / 3.0)
        m.qq.load_ijk(problems.turbulence(g, device="cpu"))
        fx = None
        if ybc == "fixed":
            fx = mhd.Fields(g)
            fx.data.copy_(m.qq.data)
        per = {"periodic": (1, 1, 1), "mirror": (1, 0, 1), "fixed": (1, 2, 1)}[ybc]
        bc = mhd.BoundaryCondition(per, sym if ybc == "mirror" else None, fixed=fx)
        m.integrator.apply_boundary_condition(bc, m.qq)
        start = m.qq.data.clone()
        outs = []
        for tiles in ((2, 8, 64, 0, 32, 0), (2, 8, 64, 0, 32, 1, 64), (2, 8, 4, 0, 32, 1, 4),
                      (1, 1, 7, 5, 16, 1, 5), (1, 1, 256, 5, 256, 1, 256),
                      (1, 1, 7, 5, 16, 2, 5), (1, 1, 256, 4, 256, 2, 256),
                      (1, 1, 7, 6, 16, 1, 5), (1, 1, 256, 6, 5, 2, 256),
                      (1, 1, 7, 5, 16, 3, 5), (1, 1, 4, 6, 16, 4, 4)):
            m.qq.data.copy_(start)
            set_tiles(ctx, *tiles)
            dts = m.integrator.run(bc, m.qq, 2)
            outs.append((m.qq.data.clone(), dts))
        set_tiles(ctx, *default)
        for n, (o, d) in enumerate(outs[1:], 1):
            assert d == outs[0][1], (sizes, n)
            assert torch.equal(o, outs[0][0]), (sizes, n, (o - outs[0][0]).abs().max().item())


def test_timed_rounds_cancel_a_steady_clock_drift():
    """Two passes, the second in reverse order, each candidate's mean: a
    clock that slows every later timing by the same step adds the same to
    every candidate, so the ranking is the candidates' own; one pass would
    have ranked the first-timed candidate fastest."""
    from miso_amd.autotune import _timed_rounds
    order = []
    own = {3: 1.03, 1: 1.01, 2: 1.02, 0: 1.00}  # candidate 0 is the fastest

    def time_one(c):
        order.append(c)
        return own[c] + 0.05 * (len(order) - 1)   # 5 % slower per timing
    cands = [3, 1, 2, 0]
    got = _timed_rounds(cands, time_one)
    assert order == cands + cands[::-1]
    assert min(got, key=got.get) == 0
    assert sorted(got, key=got.get) == [0, 1, 2, 3]
    drift = {c: got[c] - own[c] for c in cands}
    assert max(drift.values()) - min(drift.values()) < 1e-12
    one_pass = {c: own[c] + 0.05 * n for n, c in enumerate(cands)}
    assert min(one_pass, key=one_pass.get) == 3
