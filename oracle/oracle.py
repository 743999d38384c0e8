"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (mhd_oracle.c) restates MISO's ``backend::Host`` MHD step in C.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module; the product (``miso_amd``) never does.

Arrays use MISO's host order: a field set is a numpy array of shape
``(9, i_total, j_total, k_total)`` (k fastest), fields ordered
ro, vx, vy, vz, bx, by, bz, ei, ph (miso/include/miso/mhd_fields.hpp:13-34).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmiso_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libmiso_ref_pins.so")


class OgGrid(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "i_size", "j_size", "k_size", "i_total", "j_total", "k_total",
        "i_margin", "j_margin", "k_margin", "is_", "js", "ks")]


class OgParams(C.Structure):
    _fields_ = [
        ("gm", C.c_double), ("cfl_number", C.c_double), ("ep", C.c_double),
        ("fh", C.c_double), ("cs_fac", C.c_double), ("ca_fac", C.c_double),
        ("vv_fac", C.c_double), ("periodic", C.c_int * 3),
        ("sym_sign", (C.c_int * 9) * 3), ("src_kind", C.c_int),
        ("grav_axis", C.c_int), ("grav", C.c_double), ("src_heat", C.c_double),
        ("fixed", C.c_void_p)]


def build(force: bool = False) -> str:
    """Compile the oracle with the flags its parity relies on."""
    src = os.path.join(HERE, "mhd_oracle.c")
    deps = [src, os.path.join(HERE, "mhd_oracle_body.inc"), os.path.join(HERE, "mhd_oracle.h"),
            os.path.join(HERE, "rt_oracle.c")]
    if (not force and os.path.exists(LIB_PATH)
            and all(os.path.getmtime(LIB_PATH) >= os.path.getmtime(d) for d in deps)):
        return LIB_PATH
    subprocess.run(["make", "-C", HERE, "-B" if force else "-s", "oracle"], check=True)
    return LIB_PATH


def build_ref() -> str | None:
    """Compile oracle/_ref from the reference's own headers when present."""
    if not os.path.isdir("/root/reference/miso/include"):
        return None
    subprocess.run(["make", "-C", HERE, "-B", "ref"], check=True)
    return REF_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for S, R in (("f64", C.c_double), ("f32", C.c_float)):
            P = C.POINTER(R)
            getattr(L, f"og_grid_build_{S}").argtypes = [
                C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 6 + [
                C.POINTER(OgGrid), P, P]
            getattr(L, f"og_grid_build_{S}").restype = C.c_int
            getattr(L, f"og_ghost_fill_{S}").argtypes = [C.POINTER(OgGrid), C.POINTER(OgParams), P]
            getattr(L, f"og_rk_stage_{S}").argtypes = [
                C.POINTER(OgGrid), P, C.POINTER(OgParams), R, R, R, P, P, P, P]
            getattr(L, f"og_char_vel_{S}").argtypes = [
                C.POINTER(OgGrid), C.POINTER(OgParams), P, P, P]
            getattr(L, f"og_av_sweep_{S}").argtypes = [
                C.POINTER(OgGrid), P, C.POINTER(OgParams), C.c_int, R, P, P, P]
            getattr(L, f"og_cfl_{S}").argtypes = [C.POINTER(OgGrid), P, C.POINTER(OgParams), P, P]
            getattr(L, f"og_cfl_{S}").restype = R
            getattr(L, f"og_update_{S}").argtypes = [
                C.POINTER(OgGrid), P, R, C.POINTER(OgParams), R, P, P]
            getattr(L, f"og_run_{S}").argtypes = [
                C.POINTER(OgGrid), P, R, C.POINTER(OgParams), C.c_int, P, P,
                C.POINTER(C.c_double)]
            getattr(L, f"og_dqq_eval_{S}").argtypes = [R, R, R, R]
            getattr(L, f"og_dqq_eval_{S}").restype = R
            getattr(L, f"og_flux_core_{S}").argtypes = [R, R, R, R, R, R]
            getattr(L, f"og_flux_core_{S}").restype = R
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int)
        L.og_rt_quadrature.argtypes = [C.c_int, D, D, D, D]
        L.og_rt_quadrature.restype = C.c_int
        L.og_rt_solve.argtypes = [C.POINTER(OgGrid), D, C.c_int, D, D, D, I, I, D, C.c_double,
                                  C.c_int, D]
        L.og_rt_solve.restype = C.c_int
        L.og_rt_qrad.argtypes = [C.POINTER(OgGrid), C.c_int, D, D, D, D]
        L.og_work_elems.argtypes = [C.POINTER(OgGrid)]
        L.og_work_elems.restype = C.c_size_t
        L.og_set_threads.argtypes = [C.c_int]
        L.og_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _sfx(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _ptr(a: np.ndarray):
    R = C.c_double if a.dtype == np.float64 else C.c_float
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(R))


@dataclass
class Grid:
    """A MISO grid built the way grid.hpp:22-99,213-242 builds it."""
    g: OgGrid
    metric: np.ndarray
    min_dxyz: float
    dtype: np.dtype

    @property
    def shape(self):
        return (self.g.i_total, self.g.j_total, self.g.k_total)

    def axis(self, name: str) -> np.ndarray:
        It, Jt, Kt = self.shape
        offs = {"x": 0, "y": It, "z": It + Jt}
        lens = {"x": It, "y": Jt, "z": Kt}
        base = {"": 0, "d": It + Jt + Kt, "di": 2 * (It + Jt + Kt)}
        if len(name) == 1:
            pre, ax = "", name
        elif len(name) == 2:
            pre, ax = "d", name[1]
        else:
            pre, ax = "di", name[1]
        o = base[pre] + offs[ax]
        return self.metric[o:o + lens[ax]]


def make_grid(sizes, margin=2, bounds=((0.0, 1.0),) * 3, dtype=np.float64) -> Grid:
    L = lib()
    g = OgGrid()
    S = _sfx(dtype)
    rc = getattr(L, f"og_grid_build_{S}")(sizes[0], sizes[1], sizes[2], margin,
                                          bounds[0][0], bounds[0][1], bounds[1][0],
                                          bounds[1][1], bounds[2][0], bounds[2][1],
                                          C.byref(g), None, None)
    if rc:
        raise ValueError("invalid grid")
    n = 3 * (g.i_total + g.j_total + g.k_total)
    metric = np.zeros(n, dtype=dtype)
    R = C.c_double if np.dtype(dtype) == np.float64 else C.c_float
    mind = R(0)
    getattr(L, f"og_grid_build_{S}")(sizes[0], sizes[1], sizes[2], margin,
                                     bounds[0][0], bounds[0][1], bounds[1][0], bounds[1][1],
                                     bounds[2][0], bounds[2][1], C.byref(g), _ptr(metric),
                                     C.byref(mind))
    return Grid(g, metric, float(mind.value), np.dtype(dtype))


@dataclass
class Params:
    gm: float = 5.0 / 3.0
    cfl_number: float = 0.5
    ep: float = 1.0
    fh: float = 1.0
    cs_fac: float = 1.0
    ca_fac: float = 1.0
    vv_fac: float = 1.0
    periodic: tuple = (1, 1, 1)
    sym_sign: tuple = field(default_factory=lambda: ((1,) * 9,) * 3)
    src_kind: int = 0
    grav_axis: int = 2
    grav: float = 0.0
    src_heat: float = 0.0  # src_kind 2: src.ei = src_heat * ro
    fixed: object = None  # (9, It, Jt, Kt) array, MISO order, for axes with periodic == 2

    def c(self) -> OgParams:
        p = OgParams()
        for n in ("gm", "cfl_number", "ep", "fh", "cs_fac", "ca_fac", "vv_fac", "grav",
                  "src_heat"):
            setattr(p, n, float(getattr(self, n)))
        for a in range(3):
            p.periodic[a] = int(self.periodic[a])
            for f in range(9):
                p.sym_sign[a][f] = int(self.sym_sign[a][f])
        p.src_kind = int(self.src_kind)
        p.grav_axis = int(self.grav_axis)
        if any(int(x) == 2 for x in self.periodic):
            assert self.fixed is not None and self.fixed.flags["C_CONTIGUOUS"], \
                "a fixed axis needs the fixed-value fields"
            p.fixed = self.fixed.ctypes.data
        return p


def _R(dtype, v):
    return (C.c_double if np.dtype(dtype) == np.float64 else C.c_float)(v)


def ghost_fill(grid: Grid, params: Params, q: np.ndarray) -> None:
    getattr(lib(), f"og_ghost_fill_{_sfx(grid.dtype)}")(C.byref(grid.g), C.byref(params.c()),
                                                        _ptr(q))


def rk_stage(grid: Grid, params: Params, dt, ch2, tau, orgn, argm, rslt) -> None:
    N = int(np.prod(grid.shape))
    work = np.zeros(4 * N, dtype=grid.dtype)
    getattr(lib(), f"og_rk_stage_{_sfx(grid.dtype)}")(
        C.byref(grid.g), _ptr(grid.metric), C.byref(params.c()), _R(grid.dtype, dt),
        _R(grid.dtype, ch2), _R(grid.dtype, tau), _ptr(orgn), _ptr(argm), _ptr(rslt), _ptr(work))


def char_vel(grid: Grid, params: Params, q) -> np.ndarray:
    cc = np.zeros(grid.shape, dtype=grid.dtype)
    cs = np.zeros(grid.shape, dtype=grid.dtype)
    getattr(lib(), f"og_char_vel_{_sfx(grid.dtype)}")(C.byref(grid.g), C.byref(params.c()),
                                                      _ptr(q), _ptr(cc), _ptr(cs))
    return cc


def av_sweep(grid: Grid, params: Params, direction: int, dt, q, rslt, cc) -> None:
    getattr(lib(), f"og_av_sweep_{_sfx(grid.dtype)}")(
        C.byref(grid.g), _ptr(grid.metric), C.byref(params.c()), int(direction),
        _R(grid.dtype, dt), _ptr(q), _ptr(rslt), _ptr(cc))


def cfl(grid: Grid, params: Params, q) -> float:
    cs = np.zeros(grid.shape, dtype=grid.dtype)
    return float(getattr(lib(), f"og_cfl_{_sfx(grid.dtype)}")(
        C.byref(grid.g), _ptr(grid.metric), C.byref(params.c()), _ptr(q), _ptr(cs)))


def _work(grid: Grid) -> np.ndarray:
    return np.zeros(lib().og_work_elems(C.byref(grid.g)), dtype=grid.dtype)


def update(grid: Grid, params: Params, dt, q) -> None:
    getattr(lib(), f"og_update_{_sfx(grid.dtype)}")(
        C.byref(grid.g), _ptr(grid.metric), _R(grid.dtype, grid.min_dxyz), C.byref(params.c()),
        _R(grid.dtype, dt), _ptr(q), _ptr(_work(grid)))


def run(grid: Grid, params: Params, q, nsteps: int) -> np.ndarray:
    """nsteps of ModelBase::update (cfl then update); returns the dts."""
    dts = np.zeros(max(nsteps, 1), dtype=np.float64)
    getattr(lib(), f"og_run_{_sfx(grid.dtype)}")(
        C.byref(grid.g), _ptr(grid.metric), _R(grid.dtype, grid.min_dxyz), C.byref(params.c()),
        int(nsteps), _ptr(q), _ptr(_work(grid)), dts.ctypes.data_as(C.POINTER(C.c_double)))
    return dts[:nsteps]


def dqq_eval(a, b, c, ep, dtype=np.float64):
    R = _sfx(dtype)
    return getattr(lib(), f"og_dqq_eval_{R}")(a, b, c, ep)


def flux_core(qq_dw, qq_up, dqq_dw, dqq_up, cc, fh, dtype=np.float64):
    R = _sfx(dtype)
    return getattr(lib(), f"og_flux_core_{R}")(qq_dw, qq_up, dqq_dw, dqq_up, cc, fh)


def set_threads(n: int) -> None:
    lib().og_set_threads(int(n))


def get_threads() -> int:
    return int(lib().og_get_threads())


# ---------------------------------------------------------------- radiative transfer
def rt_quadrature(num_rays: int = 24):
    """(weights, mu_x, mu_y, mu_z) of rt_quadrature.hpp."""
    arrs = [np.zeros(24) for _ in range(4)]
    n = lib().og_rt_quadrature(num_rays, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    return tuple(a[:n].copy() for a in arrs)


def rt_solve(grid: Grid, kappa, S, rint, periodic=(1, 1, 0), face_kind=(0, 0, 0, 0, 1, 2),
             face_value=None, tolerance=1e-8, max_iters=200, num_rays=24):
    """RT::solve on one rank (FP64); rint (rays, It, Jt, Kt) is updated in
    place.  face_kind per (axis, side): 0 none, 1 I=S, 2 I=0, 3 face_value.
    Returns (iterations, last max difference)."""
    assert grid.dtype == np.float64
    dxyz = np.array([grid.axis("dx")[0], grid.axis("dy")[0], grid.axis("dz")[0]])
    per = (C.c_int * 3)(*periodic)
    fk = (C.c_int * 6)(*face_kind)
    fv = (C.c_double * 6)(*(face_value or (0.0,) * 6))
    md = C.c_double()
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    it = lib().og_rt_solve(C.byref(grid.g), P(dxyz), num_rays, P(kappa), P(S), P(rint), per, fk,
                           fv, tolerance, max_iters, C.byref(md))
    return it, md.value


def rt_qrad(grid: Grid, kappa, S, rint, num_rays=24) -> np.ndarray:
    q = np.zeros(grid.shape)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    lib().og_rt_qrad(C.byref(grid.g), num_rays, P(kappa), P(S), P(rint), P(q))
    return q
