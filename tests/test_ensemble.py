"""Batched ensemble (misohip_batch_*): every run of a batch equals the same
run stepped alone through misohip_step, bit for bit, in direct launches and
in captured-graph replays; FP64 runs equal the oracle."""
import numpy as np
import pytest

from miso_amd.ensemble import kh_sizes

# x extents off and on the y walk's 60-output strips (61: a strip with one
# output), y extents off and on its 16-row chunks (33, 65: a one-row chunk)
SHAPES = [(32, 32, 1), (48, 40, 1), (64, 64, 1), (40, 72, 1), (96, 32, 1), (61, 33, 1),
          (120, 65, 1)]


def test_kh_sizes_are_seeded_draws():
    a, b = kh_sizes(64, 7), kh_sizes(64, 7)
    assert a == b and set(a) <= {128, 256, 512, 1024} and len(set(a)) > 1

def _initial(seed):
    from miso_amd import problems

    def f(m, g):
        return problems.kelvin_helmholtz(g, seed=seed + m, device="cpu")
    return f


def _single(ctx, shape, init, m, dtype, nsteps):
    from miso_amd.ensemble import KH_BC, KH_BOUNDS
    from miso_amd.mhd import MHD, Config, Grid
    g = Grid(*shape, 2, bounds=KH_BOUNDS, dtype=dtype, device="cuda:0")
    model = MHD(Config({"mhd": {"cfl_number": 0.4}}), g, ctx, 1.4)
    model.qq.load_ijk(init(m, g))
    model.integrator.apply_boundary_condition(KH_BC, model.qq)
    dts = [model.integrator.step(KH_BC, model.qq) for _ in range(nsteps)]
    return model.qq.data.clone(), dts


def _interior(t, g):
    return t[:, g.k_margin:g.k_total - g.k_margin, g.j_margin:g.j_total - g.j_margin,
             g.i_margin:g.i_total - g.i_margin]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_equals_single_runs(ctx, dtype):
    import torch
    from miso_amd.ensemble import Ensemble
    init = _initial(5)
    n = 3
    ens = Ensemble(ctx, SHAPES, init, dtype=dtype)
    ens.step(n)
    dt, time, steps, fb = ens.state()
    assert fb == 0 and list(steps) == [n] * len(SHAPES)
    for m, shape in enumerate(SHAPES):
        ref, dts = _single(ctx, shape, init, m, dtype, n)
        g = ens.runs[m].grid
        got = ens.runs[m].qq.data
        assert dt[m] == dts[-1], (m, dt[m], dts)
        assert time[m] == sum(dts)
        a, b = _interior(got, g), _interior(ref, g)
        assert torch.equal(a, b), (m, (a - b).abs().max().item())
    ens.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32])
def test_graph_replays_equal_direct_launches(ctx, dtype):
    import torch
    from miso_amd.ensemble import Ensemble
    init = _initial(9)
    direct = Ensemble(ctx, SHAPES, init, dtype=dtype)
    graph = Ensemble(ctx, SHAPES, init, dtype=dtype)
    graph.capture(2)
    for rep in range(3):
        direct.step(2)
        graph.step(2)
        torch.cuda.synchronize()
        d1, t1, s1, f1 = direct.state()
        d2, t2, s2, f2 = graph.state()
        assert np.array_equal(d1, d2) and np.array_equal(t1, t2) and np.array_equal(s1, s2)
        assert f1 == 0 and f2 == 0
        for m in range(len(SHAPES)):
            assert torch.equal(direct.runs[m].qq.data, graph.runs[m].qq.data), (rep, m)
    direct.close()
    graph.close()


@pytest.mark.gpu
def test_batch_fp64_equals_oracle(ctx):
    from miso_amd import problems
    from miso_amd.ensemble import KH_BOUNDS, Ensemble
    from oracle import oracle as O
    from tests.parity_util import assert_bitwise
    shapes = [(32, 24, 1), (40, 40, 1)]
    init = _initial(2)
    ens = Ensemble(ctx, shapes, init, dtype=np.float64)
    ens.step(4)
    dt, _, _, _ = ens.state()
    sym = ((1,) * 9, (1, 1, -1, 1, 1, -1, 1, 1, 1), (1,) * 9)
    for m, shape in enumerate(shapes):
        og = O.make_grid(shape, 2, bounds=KH_BOUNDS)
        g = ens.runs[m].grid
        q = problems.to_numpy(problems.kelvin_helmholtz(g, seed=2 + m, device="cpu"))
        p = O.Params(cfl_number=0.4, gm=1.4, periodic=(1, 0, 1), sym_sign=sym)
        O.ghost_fill(og, p, q)
        dts = O.run(og, p, q, 4)
        assert dt[m] == dts[-1]
        out = ens.runs[m].qq.to_host(ctx)
        sl = (slice(None),) + g.interior()
        assert_bitwise(out[sl], q[sl], f"batch run {m}")
    ens.close()
