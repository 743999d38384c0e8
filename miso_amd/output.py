"""MISO's output files, written and read on the host, so GPU runs load in
pymiso and MISO's own tools: the layout of one run directory is

    <save_dir>/grid.bin                      Grid::save        (grid.hpp:384-422)
    <save_dir>/time/time.<n>.txt             Time::save        (time.hpp:68-88)
    <save_dir>/time/n_output.txt
    <save_dir>/mpi/coords.csv                mpi::Shape::save  (mpi_util.hpp:111-126)
    <save_dir>/mhd/mhd.<n>.<rank>.bin        Checkpoint::save  (mhd_checkpoint.hpp:29-61)
    <save_dir>/rt/rank_<rank>.bin            RT::save          (rt_host.hpp:173-199, the
                                             searchlight demo's name, main.cpp:117-122)
    <save_dir>/config.yaml                   Config::save      (config.hpp:127-142)

with <n> and <rank> zero-padded to `n_output_digits` (util::zfill,
utility.hpp:62-66) and the directory names of MISO's io defaults
(config_defaults.hpp:4-17).  Binary files are little-endian: a uint32
element size (4 or 8), then the arrays in MISO's memory order -- the grid
axes x, y, z with their ghost cells, and the nine fields ro vx vy vz bx by
bz ei ph, each (i_total, j_total, k_total) with k fastest (array3d.hpp:54-60).
The RT file has no element-size header: an int32 ray count, the weights and
the three direction cosines of every ray, then S, kappa and the intensities
(num_rays, i_total, j_total, k_total), all in the solver's element type.
A run is one rank on one GPU, so coords.csv holds rank 0 at (0, 0, 0).
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

FIELD_NAMES = ("ro", "vx", "vy", "vz", "bx", "by", "bz", "ei", "ph")
IO_DEFAULTS = {"save_dir": "data/", "time_save_dir": "time/", "mpi_save_dir": "mpi/",
               "mhd_save_dir": "mhd/", "rt_save_dir": "rt/", "n_output_digits": 8}


def zfill(num: int, width: int) -> str:
    """util::zfill: the integer padded with zeros to `width` characters."""
    return str(int(num)).rjust(width, "0")


def _elem(dtype) -> int:
    size = np.dtype(dtype).itemsize
    if size not in (4, 8):
        raise ValueError("MISO files hold float32 or float64 arrays")
    return size


def _ftype(size: int) -> str:
    return "<f4" if size == 4 else "<f8"


@dataclass
class Time:
    """Time<Real> (time.hpp:16-114): simulation time, output counter, step count."""
    time: float = 0.0
    n_output: int = 0
    n_step: int = 0

    def update(self, dt: float) -> None:
        self.time += dt
        self.n_step += 1


class RunFiles:
    """The output files of one run directory.  `io` holds MISO's io keys
    (save_dir, time_save_dir, mpi_save_dir, mhd_save_dir, n_output_digits);
    save_dir is taken relative to `config_dir`, as Config does
    (config.hpp:114-115)."""

    def __init__(self, config_dir: str, io: dict | None = None, rank: int = 0):
        self.io = dict(IO_DEFAULTS, **(io or {}))
        self.save_dir = os.path.join(config_dir, self.io["save_dir"])
        self.digits = int(self.io["n_output_digits"])
        self.rank = int(rank)

    # -- paths -----------------------------------------------------------------
    def _dir(self, key: str) -> str:
        return os.path.join(self.save_dir, self.io[key])

    def mhd_filename(self, n_output: int) -> str:
        return os.path.join(self._dir("mhd_save_dir"),
                            f"mhd.{zfill(n_output, self.digits)}.{zfill(self.rank, self.digits)}.bin")

    def time_filename(self, n_output: int) -> str:
        return os.path.join(self._dir("time_save_dir"), f"time.{zfill(n_output, self.digits)}.txt")

    # -- grid --------------------------------------------------------------------
    def save_grid(self, grid) -> None:
        """grid.bin: element size, then x, y, z with their ghost cells."""
        os.makedirs(self.save_dir, exist_ok=True)
        size = _elem(grid.dtype)
        with open(os.path.join(self.save_dir, "grid.bin"), "wb") as f:
            f.write(struct.pack("<I", size))
            for s in (grid.x, grid.y, grid.z):
                np.asarray(s).astype(_ftype(size)).tofile(f)

    def load_grid(self, totals):
        """(x, y, z) from grid.bin, given (i_total, j_total, k_total)."""
        with open(os.path.join(self.save_dir, "grid.bin"), "rb") as f:
            (size,) = struct.unpack("<I", f.read(4))
            return tuple(np.fromfile(f, dtype=_ftype(size), count=n) for n in totals)

    # -- MPI shape -----------------------------------------------------------------
    def save_coords(self) -> None:
        d = self._dir("mpi_save_dir")
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "coords.csv"), "w") as f:
            f.write("rank,x,y,z\n")
            f.write(f"{self.rank},0,0,0\n")

    # -- time ------------------------------------------------------------------------
    def save_time(self, t: Time) -> None:
        d = self._dir("time_save_dir")
        os.makedirs(d, exist_ok=True)
        with open(self.time_filename(t.n_output), "w") as f:
            f.write(f"{_fmt_real(t.time)}\n{t.n_output}\n{t.n_step}\n")
        with open(os.path.join(d, "n_output.txt"), "w") as f:
            f.write(f"{t.n_output}\n")

    def load_time(self) -> Time:
        with open(os.path.join(self._dir("time_save_dir"), "n_output.txt")) as f:
            n = int(f.read().split()[0])
        with open(self.time_filename(n)) as f:
            tok = f.read().split()
        return Time(float(tok[0]), int(tok[1]), int(tok[2]))

    # -- MHD fields ------------------------------------------------------------------
    def save_mhd(self, t: Time, q_host: np.ndarray) -> str:
        """Checkpoint::save of (9, It, Jt, Kt) host fields in MISO order."""
        if q_host.ndim != 4 or q_host.shape[0] != 9:
            raise ValueError("expected the nine fields as (9, i_total, j_total, k_total)")
        size = _elem(q_host.dtype)
        os.makedirs(self._dir("mhd_save_dir"), exist_ok=True)
        name = self.mhd_filename(t.n_output)
        with open(name, "wb") as f:
            f.write(struct.pack("<I", size))
            for m in range(9):
                np.ascontiguousarray(q_host[m]).astype(_ftype(size)).tofile(f)
        return name

    def load_mhd(self, n_output: int, shape, dtype) -> np.ndarray:
        """Checkpoint::load: the element size must match `dtype`
        (mhd_checkpoint.hpp:73-76)."""
        size = _elem(dtype)
        with open(self.mhd_filename(n_output), "rb") as f:
            (stored,) = struct.unpack("<I", f.read(4))
            if stored != size:
                raise RuntimeError("Checkpoint file element size does not match Real type size.")
            n = int(np.prod(shape))
            out = np.fromfile(f, dtype=_ftype(size), count=9 * n)
        if out.size != 9 * n:
            raise RuntimeError("Checkpoint file is shorter than nine fields of this grid")
        return out.reshape((9,) + tuple(shape)).astype(dtype)

    # -- RT ----------------------------------------------------------------------
    def rt_filename(self) -> str:
        return os.path.join(self._dir("rt_save_dir"), f"rank_{zfill(self.rank, self.digits)}.bin")

    def save_rt(self, rt) -> str:
        """RT::save of a `miso_amd.rt.RT` (its quadrature, S, kappa and
        intensities, moved to MISO order on the host)."""
        w, mx, my, mz = rt.quadrature()
        n = len(w)
        ft = _ftype(_elem(rt.grid.dtype))
        os.makedirs(self._dir("rt_save_dir"), exist_ok=True)
        name = self.rt_filename()
        with open(name, "wb") as f:
            f.write(struct.pack("<i", n))
            for a in (w, mx, my, mz):
                np.asarray(a).astype(ft).tofile(f)
            for t in (rt.src_func, rt.abs_coeff):
                np.ascontiguousarray(t.permute(2, 1, 0).cpu().numpy()).astype(ft).tofile(f)
            rint = rt.intensities()[:n]
            np.ascontiguousarray(rint.permute(0, 3, 2, 1).cpu().numpy()).astype(ft).tofile(f)
        return name

    def load_rt(self, shape, dtype=np.float64) -> dict:
        """The arrays of an RT file, given the grid's (i_total, j_total, k_total)."""
        ft = _ftype(_elem(dtype))
        cells = int(np.prod(shape))
        with open(self.rt_filename(), "rb") as f:
            (n,) = struct.unpack("<i", f.read(4))
            w, mx, my, mz = (np.fromfile(f, dtype=ft, count=n) for _ in range(4))
            src = np.fromfile(f, dtype=ft, count=cells).reshape(shape)
            kap = np.fromfile(f, dtype=ft, count=cells).reshape(shape)
            rint = np.fromfile(f, dtype=ft, count=n * cells).reshape((n,) + tuple(shape))
        return {"num_rays": n, "weights": w, "mu_x": mx, "mu_y": my, "mu_z": mz,
                "src_func": src, "abs_coeff": kap, "rint": rint}

    # -- config --------------------------------------------------------------------
    def save_config(self, config: dict) -> None:
        """config.yaml in save_dir (YAML of the run's configuration, with the
        io keys this writer used)."""
        import yaml
        os.makedirs(self.save_dir, exist_ok=True)
        c = dict(config)
        c["io"] = dict(c.get("io", {}), **self.io)
        with open(os.path.join(self.save_dir, "config.yaml"), "w") as f:
            yaml.safe_dump(c, f, sort_keys=False)

    def save_run(self, grid, t: Time, q_host: np.ndarray) -> str:
        """Everything a MISO run leaves for one output: grid, coords, time, fields."""
        self.save_grid(grid)
        self.save_coords()
        self.save_time(t)
        return self.save_mhd(t, q_host)


def _fmt_real(x: float) -> str:
    """A time value as `std::ostream << Real` writes it (6 significant digits,
    %g style)."""
    return f"{x:g}"


__all__ = ["RunFiles", "Time", "FIELD_NAMES", "IO_DEFAULTS", "zfill"]
