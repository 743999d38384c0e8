"""MISO's output files (miso_amd/output.py): byte layout of grid.bin and the
MHD checkpoint (grid.hpp:384-422, mhd_checkpoint.hpp:29-61 -- the same
layout pymiso's fixtures write, pymiso/tests/test_pymiso_data.py:14-46),
the time and coords text files (time.hpp:68-88, mpi_util.hpp:111-126), file
naming (util::zfill), round trips and the element-size check."""
import os
import struct

import numpy as np
import pytest

from miso_amd.mhd import Grid
from miso_amd.output import RunFiles, Time, zfill


def test_zfill_pads_like_the_reference():
    assert zfill(3, 8) == "00000003"
    assert zfill(12345678, 8) == "12345678"
    assert zfill(0, 4) == "0000"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_run_files_have_miso_layout(tmp_path, dtype):
    g = Grid(6, 4, 1, 2, bounds=((0.0, 3.0), (-1.0, 1.0), (0.0, 1.0)), dtype=dtype, device="cpu")
    rng = np.random.default_rng(1)
    q = rng.standard_normal((9,) + g.shape).astype(dtype)
    rf = RunFiles(str(tmp_path), rank=0)
    name = rf.save_run(g, Time(1.5, 3, 7), q)
    base = tmp_path / "data"
    assert name == str(base / "mhd" / "mhd.00000003.00000000.bin")
    size = np.dtype(dtype).itemsize
    ft = "<f4" if size == 4 else "<f8"
    # grid.bin: element size, then the axes with ghosts (x: 6 + 4, y: 4 + 4, z: 1)
    raw = (base / "grid.bin").read_bytes()
    assert raw[:4] == struct.pack("<I", size)
    x = np.frombuffer(raw[4:], dtype=ft)
    assert x.size == 10 + 8 + 1
    dx = 3.0 / 6
    np.testing.assert_allclose(x[:10], 0.0 + (np.arange(10) - 2 + 0.5) * dx, rtol=1e-6)
    np.testing.assert_allclose(x[10:18], -1.0 + (np.arange(8) - 2 + 0.5) * 0.5, rtol=1e-6)
    assert x[18] == np.asarray(0.5, dtype=ft)
    # checkpoint: element size, then nine (It, Jt, Kt) arrays, k fastest
    raw = open(name, "rb").read()
    expect = struct.pack("<I", size) + b"".join(q[m].astype(ft).ravel(order="C").tobytes()
                                               for m in range(9))
    assert raw == expect
    assert (base / "time" / "time.00000003.txt").read_text() == "1.5\n3\n7\n"
    assert (base / "time" / "n_output.txt").read_text() == "3\n"
    assert (base / "mpi" / "coords.csv").read_text() == "rank,x,y,z\n0,0,0,0\n"


def test_grid_bin_matches_the_pymiso_fixture(tmp_path):
    """The grid of pymiso's fixture (4 x 2 x 1 cells, margin 1, x in [0, 4],
    y in [0, 2]; test_pymiso_data.py:171-213) written by this writer gives
    the fixture's bytes."""
    g = Grid(4, 2, 1, 1, bounds=((0.0, 4.0), (0.0, 2.0), (0.0, 1.0)), dtype=np.float32,
             device="cpu")
    RunFiles(str(tmp_path)).save_grid(g)
    expect = struct.pack("<I", 4) + b"".join(
        np.array(v, dtype="<f4").tobytes()
        for v in ([-0.5, 0.5, 1.5, 2.5, 3.5, 4.5], [-0.5, 0.5, 1.5, 2.5], [0.5]))
    assert (tmp_path / "data" / "grid.bin").read_bytes() == expect


def test_round_trip_and_element_size_check(tmp_path):
    g = Grid(5, 3, 2, 2, dtype=np.float64, device="cpu")
    q = np.arange(9 * int(np.prod(g.shape)), dtype=np.float64).reshape((9,) + g.shape)
    rf = RunFiles(str(tmp_path), io={"save_dir": "out/", "n_output_digits": 5})
    rf.save_run(g, Time(0.25, 12, 40), q)
    assert os.path.exists(tmp_path / "out" / "mhd" / "mhd.00012.00000.bin")
    assert np.array_equal(rf.load_mhd(12, g.shape, np.float64), q)
    t = rf.load_time()
    assert (t.time, t.n_output, t.n_step) == (0.25, 12, 40)
    xs = rf.load_grid((g.i_total, g.j_total, g.k_total))
    for got, ref in zip(xs, (g.x, g.y, g.z)):
        assert np.array_equal(got, ref)
    with pytest.raises(RuntimeError, match="element size"):
        rf.load_mhd(12, g.shape, np.float32)


def test_time_text_uses_six_significant_digits(tmp_path):
    rf = RunFiles(str(tmp_path))
    rf.save_time(Time(0.000123456789, 1, 2))
    assert (tmp_path / "data" / "time" / "time.00000001.txt").read_text() == "0.000123457\n1\n2\n"
    rf.save_time(Time(1234567.0, 2, 3))
    assert (tmp_path / "data" / "time" / "time.00000002.txt").read_text() == "1.23457e+06\n2\n3\n"


def test_rt_file_has_miso_layout(tmp_path):
    """RT::save layout (rt_host.hpp:173-199): int32 ray count, weights and
    direction cosines (the reference's own quadrature, committed vectors),
    S, kappa and the intensities in MISO order; read back equal."""
    import os.path as op

    import torch

    from miso_amd.rt import RT
    g = Grid(4, 3, 2, 2, dtype=np.float64, device="cpu")
    rt = RT(g, 24)
    rng = np.random.default_rng(3)
    for t in (rt.src_func, rt.abs_coeff, rt.rint):
        t.copy_(torch.from_numpy(rng.standard_normal(tuple(t.shape))))
    rf = RunFiles(str(tmp_path))
    name = rf.save_rt(rt)
    assert name == str(tmp_path / "data" / "rt" / "rank_00000000.bin")
    ref = np.load(op.join(op.dirname(__file__), "golden", "ref_limiter_quadrature.npz"))
    q = ref["quadrature_24"]  # columns: weight, mu_x, mu_y, mu_z
    shape = (g.i_total, g.j_total, g.k_total)
    expect = (struct.pack("<i", 24) + b"".join(q[:, c].astype("<f8").tobytes() for c in range(4))
              + rt.src_func.permute(2, 1, 0).numpy().astype("<f8").tobytes()
              + rt.abs_coeff.permute(2, 1, 0).numpy().astype("<f8").tobytes()
              + rt.rint.permute(0, 3, 2, 1).numpy().astype("<f8").tobytes())
    assert open(name, "rb").read() == expect
    back = rf.load_rt(shape)
    assert back["num_rays"] == 24
    assert np.array_equal(back["rint"], rt.rint.permute(0, 3, 2, 1).numpy())


def test_config_yaml_records_the_io_keys(tmp_path):
    import yaml
    rf = RunFiles(str(tmp_path), io={"n_output_digits": 6})
    rf.save_config({"mhd": {"cfl_number": 0.4}, "grid": {"i_size": 8}})
    c = yaml.safe_load((tmp_path / "data" / "config.yaml").read_text())
    assert c["mhd"]["cfl_number"] == 0.4 and c["grid"]["i_size"] == 8
    assert c["io"]["n_output_digits"] == 6 and c["io"]["mhd_save_dir"] == "mhd/"


@pytest.mark.gpu
def test_gpu_run_checkpoint_round_trip(ctx, tmp_path):
    """Fields of a GPU run written as a MISO checkpoint and read back are the
    device fields bit for bit; loading them into a new run continues it
    identically."""
    import torch
    from miso_amd import mhd, problems
    g = mhd.Grid(24, 20, 16, 2, device="cuda:0")
    bc = mhd.BoundaryCondition((1, 1, 1))
    m = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, 5.0 / 3.0)
    m.qq.load_ijk(problems.turbulence(g, device="cpu"))
    m.integrator.apply_boundary_condition(bc, m.qq)
    t = Time()
    for dt in m.integrator.run(bc, m.qq, 3):
        t.update(dt)
    t.n_output = 1
    rf = RunFiles(str(tmp_path))
    rf.save_run(g, t, m.qq.to_host(ctx))
    q = rf.load_mhd(1, g.shape, np.float64)
    assert np.array_equal(q, m.qq.to_host(ctx))
    m2 = mhd.MHD(mhd.Config({"mhd": {"cfl_number": 0.4}}), g, ctx, 5.0 / 3.0)
    m2.apply_initial_condition(q, bc)
    d1 = m.integrator.run(bc, m.qq, 2)
    d2 = m2.integrator.run(bc, m2.qq, 2)
    assert d1 == d2
    assert torch.equal(m.qq.data, m2.qq.data)
