"""Python host mirror of MISO's MHD interface over the HIP backend.

Names and argument meanings follow the reference's C++ API so the tests read
like MISO's own:
  Config                  miso/include/miso/config.hpp:65-142 (+ defaults,
                          config_defaults.hpp:4-32)
  Grid                    miso/include/miso/grid.hpp:22-99,147-242 (uniform axes)
  Fields                  miso/include/miso/mhd_fields.hpp (device copy, x-fastest)
  Integrator              miso/include/miso/mhd_integrator_host.hpp:13-357
                          (update, cfl, apply_boundary_condition, update_sc4)
  MHD                     miso/include/miso/mhd.hpp:15-50
Device memory comes from torch (plumbing); every computation is a call into
libmisohip.so through the C ABI.
"""
from __future__ import annotations

import contextlib
import copy
import ctypes as C
import math
import traceback

import numpy as np

from . import hip

CONFIG_DEFAULTS = {
    "mhd": {
        "cfl_number": 0.5,
        "artificial_viscosity": {"ep": 1.0, "fh": 1.0, "cs_fac": 1.0, "ca_fac": 1.0,
                                 "vv_fac": 1.0},
    },
}


def _merge(target: dict, defaults: dict) -> dict:
    """merge_yaml (config.hpp:39-62): fill keys missing in target."""
    for k, v in defaults.items():
        if isinstance(v, dict):
            child = target.get(k)
            if not isinstance(child, dict):
                child = {}
            target[k] = _merge(child, v)
        elif target.get(k) is None:
            target[k] = v
    return target


class Config:
    """YAML configuration with MISO's defaults merged in."""

    def __init__(self, data: dict | str):
        if isinstance(data, str):
            import yaml
            with open(data) as f:
                data = yaml.safe_load(f)
        self.yaml_obj = _merge(copy.deepcopy(data or {}), CONFIG_DEFAULTS)

    def __getitem__(self, key):
        return self.yaml_obj[key]


def _torch():
    import torch
    return torch


def axis_grid(size: int, margin: int, vmin: float, vmax: float, dtype):
    """AxisGrid of a uniform axis (grid.hpp:22-99); returns s, ds, dsi, margin, stride."""
    R = np.dtype(dtype).type
    stride = 1 if size > 1 else 0
    m = margin * stride
    total = size + 2 * m
    min_value, max_value = R(vmin), R(vmax)
    space = R((max_value - min_value) / R(size))
    if size == 1:
        s = np.array([min_value + (R(0) + R(0.5)) * space], dtype=dtype)
        return s, np.array([R(1.e30)], dtype=dtype), np.array([R(0.0)], dtype=dtype), m, stride
    ig = np.arange(total) - m
    s = (min_value + (ig.astype(dtype) + R(0.5)) * space).astype(dtype)
    ds = np.full(total, space, dtype=dtype)
    dsi = (1.0 / ds.astype(np.float64)).astype(dtype)
    return s, ds, dsi, m, stride


def float_min3(a, b, c):
    """util::min3 -- narrows to float (utility.hpp:91-124)."""
    return float(min(np.float32(a), np.float32(b), np.float32(c)))


class Grid:
    """A one-rank MISO grid with its metric arrays mirrored on the device."""

    def __init__(self, i_size, j_size, k_size, margin=2,
                 bounds=((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), dtype=np.float64,
                 device="cuda"):
        if margin < 1 or min(i_size, j_size, k_size) <= 0:
            raise ValueError("Invalid grid size or margin")
        for lo, hi in bounds:
            if lo >= hi:
                raise ValueError("Invalid grid range")
        self.dtype = np.dtype(dtype)
        self.i_size, self.j_size, self.k_size = i_size, j_size, k_size
        self.x, self.dx, self.dxi, self.i_margin, self.is_ = axis_grid(i_size, margin, *bounds[0], dtype)
        self.y, self.dy, self.dyi, self.j_margin, self.js = axis_grid(j_size, margin, *bounds[1], dtype)
        self.z, self.dz, self.dzi, self.k_margin, self.ks = axis_grid(k_size, margin, *bounds[2], dtype)
        self.i_total, self.j_total, self.k_total = len(self.x), len(self.y), len(self.z)
        self.min_dxyz = float(self.dtype.type(float_min3(self.dx[0], self.dy[0], self.dz[0])))
        self.bounds = bounds
        self.device = device
        self._dev = None
        self._c = None

    @property
    def shape(self):
        return (self.i_total, self.j_total, self.k_total)

    @property
    def n_total(self):
        return self.i_total * self.j_total * self.k_total

    @property
    def n_interior(self):
        return self.i_size * self.j_size * self.k_size

    def interior(self):
        """Slices of the interior in MISO (i, j, k) order."""
        return (slice(self.i_margin, self.i_total - self.i_margin),
                slice(self.j_margin, self.j_total - self.j_margin),
                slice(self.k_margin, self.k_total - self.k_margin))

    def c_struct(self) -> hip.MisohipGrid:
        if self._c is None:
            torch = _torch()
            tdt = torch.float64 if self.dtype == np.float64 else torch.float32
            self._dev = {n: torch.from_numpy(np.ascontiguousarray(getattr(self, n))).to(
                device=self.device, dtype=tdt) for n in ("dx", "dy", "dz", "dxi", "dyi", "dzi")}
            g = hip.MisohipGrid()
            g.i_size, g.j_size, g.k_size = self.i_size, self.j_size, self.k_size
            g.i_total, g.j_total, g.k_total = self.shape
            g.i_margin, g.j_margin, g.k_margin = self.i_margin, self.j_margin, self.k_margin
            g.is_, g.js, g.ks = self.is_, self.js, self.ks
            g.min_dxyz = self.min_dxyz
            for n, t in self._dev.items():
                setattr(g, n, t.data_ptr())
            self._c = g
        return self._c


class Fields:
    """Nine device fields of one grid, stored x-fastest: tensor (9, Kt, Jt, It)."""

    def __init__(self, grid: Grid):
        torch = _torch()
        self.grid = grid
        tdt = torch.float64 if grid.dtype == np.float64 else torch.float32
        self.data = torch.zeros((9, grid.k_total, grid.j_total, grid.i_total), dtype=tdt,
                                device=grid.device)

    def c_struct(self) -> hip.MisohipFields:
        f = hip.MisohipFields()
        for m in range(9):
            f.q[m] = self.data[m].data_ptr()
        return f

    def copy_from_host(self, ctx, q_host: np.ndarray) -> None:
        """Fields<HIP>::copy_from(Fields<Host>); q_host is (9, It, Jt, Kt) MISO order."""
        assert q_host.shape == (9,) + self.grid.shape and q_host.dtype == self.grid.dtype
        g = self.grid.c_struct()
        dt = hip.F64 if self.grid.dtype == np.float64 else hip.F32
        for m in range(9):
            a = np.ascontiguousarray(q_host[m])
            hip.call("misohip_copy_to_device", ctx.handle, C.byref(g), dt,
                     a.ctypes.data_as(C.c_void_p), C.c_void_p(self.data[m].data_ptr()))

    def load_ijk(self, q) -> None:
        """Device-side load of a (9, It, Jt, Kt) torch tensor in MISO index
        order (a permute on the GPU; used for large initial conditions)."""
        self.data.copy_(q.to(device=self.data.device, dtype=self.data.dtype).permute(0, 3, 2, 1))

    def to_host(self, ctx) -> np.ndarray:
        """(9, It, Jt, Kt) numpy copy in MISO order."""
        out = np.empty((9,) + self.grid.shape, dtype=self.grid.dtype)
        g = self.grid.c_struct()
        dt = hip.F64 if self.grid.dtype == np.float64 else hip.F32
        for m in range(9):
            hip.call("misohip_copy_to_host", ctx.handle, C.byref(g), dt,
                     C.c_void_p(self.data[m].data_ptr()), out[m].ctypes.data_as(C.c_void_p))
        return out


class ExecContext:
    """ExecContext<Real, HIP>: device ordinal and the stream calls run on."""

    def __init__(self, device: int = 0, stream=None):
        torch = _torch()
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        h = C.c_void_p()
        hip.call("misohip_create", int(device), C.c_void_p(stream), C.byref(h))
        self.handle = h

    def synchronize(self):
        hip.call("misohip_synchronize", self.handle)

    def close(self):
        if self.handle:
            hip.lib().misohip_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BoundaryCondition:
    """Per-axis boundary: periodic self-exchange (True / 1), mirror with
    per-field signs (False / 0; boundary_condition.hpp:108-136, signs +1
    unless given), fixed ghost values ("fixed" / 2: every fill copies the
    axis' ghost cells from the `fixed` Fields -- a user boundary functor
    writing given values, the mhd_model_base.hpp:131-155 contract) or user
    ("user" / 3: `user(q)` runs before every fill -- BC::apply before the
    halo exchange, mhd_integrator_host.hpp:250-255 -- with q the (9, Kt, Jt,
    It) tensor being filled, and sets that axis' ghost cells with torch
    operations on the current stream)."""

    FIXED = 2
    USER = 3

    def __init__(self, periodic=(True, True, True), sym_sign=None, fixed=None, user=None):
        def kind(p):
            if p == "fixed" or (not isinstance(p, bool) and p == 2):
                return 2
            if p == "user" or (not isinstance(p, bool) and p == 3):
                return 3
            return int(bool(p))
        self.periodic = tuple(kind(p) for p in periodic)
        self.sym_sign = sym_sign or ((1,) * 9,) * 3
        self.fixed = fixed  # Fields holding the ghost values of fixed axes
        self.user = user
        if 2 in self.periodic and fixed is None:
            raise ValueError("a fixed boundary needs the Fields that hold its ghost values")
        if 3 in self.periodic and user is None:
            raise ValueError("a user boundary needs its function")

    def c_struct(self) -> hip.MisohipBc:
        b = hip.MisohipBc()
        for a in range(3):
            b.periodic[a] = self.periodic[a]
            for f in range(9):
                b.sym_sign[a][f] = int(self.sym_sign[a][f])
        if self.fixed is not None:
            for f in range(9):
                b.fixed[f] = self.fixed.data[f].data_ptr()
        return b


class UserSource:
    """A user source term, Source::vx, vy, vz and ei (mhd_model_base.hpp:
    131-155, e.g. demo/mhd3d_magnetosphere/src/src.hpp:4-41).  The reference
    evaluates it on argm at every interior cell of every RK stage
    (mhd_integrator_host.hpp:134,151,168,241); here fn(argm, out) runs before
    every stage with argm the (9, Kt, Jt, It) tensor of that stage and out a
    (4, Kt, Jt, It) tensor, and writes src.vx, src.vy, src.vz, src.ei of
    every interior cell into out[0..3] with torch operations on the current
    stream.  The stage adds out[0..2] to the momenta inside dt and out[3] to
    the internal energy outside it."""

    def __init__(self, fn):
        self.fn = fn


class Integrator:
    """Integrator<Real, backend::HIP> (mhd_integrator_host.hpp:13-357)."""

    def __init__(self, config: Config, grid: Grid, exec_ctx: ExecContext, gm: float,
                 src_kind: int = 0, grav_axis: int = 2, grav: float = 0.0,
                 source: UserSource | None = None):
        av = config["mhd"]["artificial_viscosity"]
        self.grid, self.ctx = grid, exec_ctx
        self.cfl_number = float(config["mhd"]["cfl_number"])
        p = hip.MisohipParams()
        p.gm = float(gm)
        p.cfl_number = self.cfl_number
        p.ep, p.fh = float(av["ep"]), float(av["fh"])
        p.cs_fac, p.ca_fac, p.vv_fac = float(av["cs_fac"]), float(av["ca_fac"]), float(av["vv_fac"])
        p.src_kind, p.grav_axis, p.grav = int(src_kind), int(grav_axis), float(grav)
        self.source = source
        if source is not None:
            p.src_kind = 2
        if p.ep < 0 or p.fh < 0:
            raise ValueError("artificial viscosity ep and fh must be >= 0")
        self.params = p
        self.qq_argm = Fields(grid)
        self.qq_rslt = Fields(grid)
        torch = _torch()
        self.cc = torch.zeros(self.qq_argm.data.shape[1:], dtype=self.qq_argm.data.dtype,
                              device=grid.device)
        self.dtype_code = hip.F64 if grid.dtype == np.float64 else hip.F32
        w = hip.MisohipWork()
        w.argm = self.qq_argm.c_struct()
        w.rslt = self.qq_rslt.c_struct()
        w.cc = self.cc.data_ptr()
        self._work = w
        self.src_out = None
        if source is not None:
            self.src_out = torch.zeros((4,) + tuple(self.qq_argm.data.shape[1:]),
                                       dtype=self.qq_argm.data.dtype, device=grid.device)

    # -- user sources ----------------------------------------------------------
    @contextlib.contextmanager
    def _user_src(self, *fields: Fields):
        """self.source as the context's source callback for one call: the
        argm a stage hands over is one of `fields` and the two workspaces."""
        if self.source is None:
            yield
            return
        tensors = {f.data[0].data_ptr(): f.data for f in fields + (self.qq_argm, self.qq_rslt)}
        out = self.src_out

        def cb(_user, q, _arrays, _stream):
            try:
                self.source.fn(tensors[q.contents.q[0]], out)
                return 0
            except Exception:  # noqa: BLE001 -- reported, then the call fails
                traceback.print_exc()
                return 1

        fn = hip.SRC_FN(cb)
        arrays = (C.c_void_p * 4)(*[out[d].data_ptr() for d in range(4)])
        hip.call("misohip_set_source_callback", self.ctx.handle, fn, None, arrays)
        try:
            yield
        finally:
            hip.lib().misohip_set_source_callback(self.ctx.handle, hip.SRC_FN(), None, None)

    # -- user boundaries -------------------------------------------------------
    @contextlib.contextmanager
    def _user_bc(self, bc: BoundaryCondition, qq: Fields):
        """bc.user as the context's boundary callback for one call: the
        fields a fill hands over are one of qq and the two workspaces; the
        user source, if any, is installed for the call as well."""
        if bc.user is None:
            with self._user_src(qq):
                yield
            return
        tensors = {f.data[0].data_ptr(): f.data for f in (qq, self.qq_argm, self.qq_rslt)}

        def cb(_user, q, _stream):
            try:
                bc.user(tensors[q.contents.q[0]])
                return 0
            except Exception:  # noqa: BLE001 -- reported, then the call fails
                traceback.print_exc()
                return 1

        fn = hip.BC_FN(cb)
        hip.call("misohip_set_bc_callback", self.ctx.handle, fn, None)
        try:
            with self._user_src(qq):
                yield
        finally:
            hip.lib().misohip_set_bc_callback(self.ctx.handle, hip.BC_FN(), None)

    # -- per-function entry points ------------------------------------------
    def apply_boundary_condition(self, bc: BoundaryCondition, qq: Fields) -> None:
        with self._user_bc(bc, qq):
            hip.call("misohip_ghost_fill", self.ctx.handle, C.byref(self.grid.c_struct()),
                     self.dtype_code, C.byref(qq.c_struct()), C.byref(bc.c_struct()))

    def divb_parameters(self, dt: float):
        """divb_parameters_set (mhd_integrator_host.hpp:337-341) in the run's Real."""
        R = self.grid.dtype.type
        dt = R(dt)
        ch = R(0.8 * float(R(self.cfl_number)) * float(R(self.grid.min_dxyz)) / float(dt))
        return R(ch * ch), R(2.0 * float(dt))

    def update_sc4(self, dt_s: float, tau: float, ch2: float, qq_orgn: Fields,
                   qq_argm: Fields, qq_rslt: Fields) -> None:
        R = self.grid.dtype.type
        damp = float(R(math.exp(float(R(-R(dt_s) / R(tau))))))
        with self._user_src(qq_orgn, qq_argm):
            hip.call("misohip_rk_stage", self.ctx.handle, C.byref(self.grid.c_struct()),
                     self.dtype_code, C.byref(self.params), C.byref(qq_orgn.c_struct()),
                     C.byref(qq_argm.c_struct()), C.byref(qq_rslt.c_struct()), float(dt_s),
                     float(ch2), damp)

    def av_sweep(self, direction: int, dt: float, qq: Fields, qq_rslt: Fields, cc) -> None:
        hip.call("misohip_av_sweep", self.ctx.handle, C.byref(self.grid.c_struct()),
                 self.dtype_code, C.byref(self.params), int(direction), float(dt),
                 C.byref(qq.c_struct()), C.byref(qq_rslt.c_struct()), C.c_void_p(cc.data_ptr()))

    # -- the step ------------------------------------------------------------
    def cfl(self, qq: Fields) -> float:
        out = C.c_double()
        hip.call("misohip_cfl", self.ctx.handle, C.byref(self.grid.c_struct()), self.dtype_code,
                 C.byref(self.params), C.byref(qq.c_struct()), C.byref(out))
        return out.value

    def update(self, dt: float, bc: BoundaryCondition, qq: Fields) -> None:
        with self._user_bc(bc, qq):
            hip.call("misohip_update", self.ctx.handle, C.byref(self.grid.c_struct()),
                     self.dtype_code, C.byref(self.params), C.byref(bc.c_struct()),
                     C.byref(qq.c_struct()), C.byref(self._work), float(dt))

    def update_heated(self, dt: float, bc: BoundaryCondition, qq: Fields, qrad) -> float:
        """update(dt), the radiative heating ei += dt * qrad / ro on the
        interior and the ghost fill (FP64); returns the next step's dt, the
        CFL of the heated state.  With x and y active the last AV sweep does
        the heating and the CFL reduction as it stores each cell: the same
        bits as update + RT.heat + apply_boundary_condition + cfl."""
        nxt = C.c_double()
        with self._user_bc(bc, qq):
            hip.call("misohip_update_heated", self.ctx.handle, C.byref(self.grid.c_struct()),
                     self.dtype_code, C.byref(self.params), C.byref(bc.c_struct()),
                     C.byref(qq.c_struct()), C.byref(self._work), float(dt),
                     C.c_void_p(qrad.data_ptr()), C.byref(nxt))
        return nxt.value

    def run(self, bc: BoundaryCondition, qq: Fields, nsteps: int):
        """nsteps of ModelBase::update (cfl + update) on qq, the CFL of each
        later step reduced by the previous step's last sweep; returns the dts."""
        dts = (C.c_double * max(1, nsteps))()
        with self._user_bc(bc, qq):
            hip.call("misohip_run", self.ctx.handle, C.byref(self.grid.c_struct()),
                     self.dtype_code, C.byref(self.params), C.byref(bc.c_struct()),
                     C.byref(qq.c_struct()), C.byref(self._work), int(nsteps), dts)
        return list(dts)[:nsteps]

    def step(self, bc: BoundaryCondition, qq: Fields) -> float:
        out = C.c_double()
        with self._user_bc(bc, qq):
            hip.call("misohip_step", self.ctx.handle, C.byref(self.grid.c_struct()),
                     self.dtype_code, C.byref(self.params), C.byref(bc.c_struct()),
                     C.byref(qq.c_struct()), C.byref(self._work), C.byref(out))
        return out.value


class MHD:
    """mhd::MHD<Real, HIP> (mhd.hpp:15-50): grid + fields + integrator."""

    def __init__(self, config: Config, grid: Grid, exec_ctx: ExecContext, gm: float, **src):
        self.grid, self.ctx = grid, exec_ctx
        self.qq = Fields(grid)
        self.integrator = Integrator(config, grid, exec_ctx, gm, **src)

    def apply_initial_condition(self, q_host: np.ndarray, bc: BoundaryCondition) -> None:
        self.qq.copy_from_host(self.ctx, q_host)
        self.integrator.apply_boundary_condition(bc, self.qq)

    def cfl(self) -> float:
        return self.integrator.cfl(self.qq)

    def update(self, dt: float, bc: BoundaryCondition) -> None:
        self.integrator.update(dt, bc, self.qq)
