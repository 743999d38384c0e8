"""Uniform-table interpolation (table_interpolator.hpp): the oracle against
the reference's compiled function (committed vectors) and the reference's
own test cases (test_table_interpolator.cpp); the device interpolators
against the same vectors bit for bit, FP64 and FP32, flat and 3D inputs."""
import os

import numpy as np
import pytest

from oracle import table_oracle as T

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_table_interp.npz")


@pytest.mark.parametrize("S", ["f64", "f32"])
def test_oracle_matches_reference_vectors(S):
    d = np.load(GOLD)
    y = T.table1d(d[f"t1_{S}"], *d[f"lohi1_{S}"], d[f"x1_{S}"])
    assert np.array_equal(y, d[f"y1_{S}"])
    y2 = T.table2d(d[f"t2_{S}"], *d[f"ranges2_{S}"], d[f"x20_{S}"], d[f"x21_{S}"])
    assert np.array_equal(y2, d[f"y2_{S}"])


def test_oracle_reference_cases():
    """The cases of test_table_interpolator.cpp:8-166 (1D: i^2 on 5 nodes;
    2D: 100 i^2 + j^2 on 2 x 3 nodes): nodes, midpoints, outside."""
    t1 = np.array([i * i for i in range(5)], np.float32)
    assert np.allclose(T.table1d(t1, 0.0, 4.0, np.arange(5, dtype=np.float32)), t1, atol=1e-5)
    mid = T.table1d(t1, 0.0, 4.0, np.arange(4, dtype=np.float32) + 0.5)
    assert np.allclose(mid, 0.5 * (t1[:-1] + t1[1:]), atol=1e-5)
    out = T.table1d(t1, 0.0, 4.0, np.array([-0.5, 4.5], np.float32))
    assert np.allclose(out, [t1[0], t1[-1]], atol=1e-5)
    t2 = np.array([[100 * i * i + j * j for j in range(3)] for i in range(2)], np.float32)
    i, j = np.meshgrid(np.arange(2), np.arange(3), indexing="ij")
    got = T.table2d(t2, 0.0, 1.0, 0.0, 2.0, i.astype(np.float32).ravel(), j.astype(np.float32).ravel())
    assert np.allclose(got, t2.ravel(), atol=1e-5)
    got = T.table2d(t2, 0.0, 1.0, 0.0, 2.0, np.full(2, 0.5, np.float32),
                    np.array([0.5, 1.5], np.float32))
    ex = [0.25 * (t2[0, j] + t2[1, j] + t2[0, j + 1] + t2[1, j + 1]) for j in range(2)]
    assert np.allclose(got, ex, atol=1e-5)
    got = T.table2d(t2, 0.0, 1.0, 0.0, 2.0, np.array([-0.5, -0.5, 1.5, 1.5], np.float32),
                    np.array([-0.5, 2.5, -0.5, 2.5], np.float32))
    assert np.allclose(got, [t2[0, 0], t2[0, 2], t2[1, 0], t2[1, 2]], atol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("S", ["f64", "f32"])
def test_device_interpolators_match_reference_vectors(ctx, S):
    import torch
    from miso_amd.table import UniformTableInterpolator1D, UniformTableInterpolator2D
    d = np.load(GOLD)
    dev = "cuda:0"
    I1 = UniformTableInterpolator1D(ctx, torch.from_numpy(d[f"t1_{S}"]).to(dev), *d[f"lohi1_{S}"])
    y = I1(torch.from_numpy(d[f"x1_{S}"]).to(dev)).cpu().numpy()
    assert np.array_equal(y, d[f"y1_{S}"])
    I2 = UniformTableInterpolator2D(ctx, torch.from_numpy(d[f"t2_{S}"]).to(dev), *d[f"ranges2_{S}"])
    x0, x1 = (torch.from_numpy(d[f"x2{k}_{S}"]).to(dev) for k in (0, 1))
    y2 = I2(x0, x1).cpu().numpy()
    assert np.array_equal(y2, d[f"y2_{S}"])
    # a 3D field, element by element (interpolate over Array3D)
    f0 = x0[:4000].reshape(10, 20, 20).contiguous()
    f1 = x1[:4000].reshape(10, 20, 20).contiguous()
    y3 = I2(f0, f1).cpu().numpy().ravel()
    assert np.array_equal(y3, d[f"y2_{S}"][:4000])
