"""Radiative MHD on one GPU: the MHD step of `mhd.py` coupled to the grey RT
solver of `rt.py`, with the per-column MLP surrogate of `surrogate.py` as an
optional replacement of most RT solves.

The reference has RT and MHD as separate models (miso/include/miso/rt_host.hpp,
mhd_integrator_host.hpp) and no coupling; this module adds one:

  step (plain)   : Q_rad from an RT solve of the current state, dt from the
                   CFL condition, the MHD update, then the operator-split
                   heating ei += dt Q_rad / ro and a ghost fill.
  step (learned) : Q_rad from the network on every step; every `rt_every`
                   steps a plain RT solve instead, whose result also sets the
                   residual indicators: flags (columns where the network
                   misses the solve by more than `flag_tol` in relative L2)
                   that take the last solved Q_rad until the next solve, and
                   the bottom slab (the optically thick layers next to the
                   bottom boundary that hold most of the network's residual
                   at the solve) where the plain solver runs on every step.

Workload: a gravity-stratified ideal gas (T falling linearly from T_bot to
T_top over the box height, hydrostatic pressure with the scale height
`h_frac` of the box height at the mean temperature), grey opacity
kappa = kappa0 ro T^3 with kappa0 set so the vertical optical depth from the
top down to mid-height is 1, S = s_rad T^4, seeded velocity and magnetic
perturbations of rms `amp`.  x and y are periodic; the z ghost cells hold
the hydrostatic profile at rest and keep it (a fixed boundary, so the
atmosphere stays supported at both ends); for RT the bottom face is
thermalised (I = S) and no light enters at the top.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .mhd import BoundaryCondition, Config, Grid, MHD
from .problems import random_modes_field, _periodic_pad
from .rt import RT

RO, VX, VY, VZ, BX, BY, BZ, EI, PH = range(9)

RT_FACES = dict(periodic=(True, True, False), face_kind=(0, 0, 0, 0, 1, 2))
# the bottom slab: thermalised at both z faces (its top face lies deep in the
# optically thick layers, where the downward intensity is the local source)
SLAB_FACES = dict(periodic=(True, True, False), face_kind=(0, 0, 0, 0, 1, 1))


class BottomSlabRT:
    """Plain RT solve of the bottom `layers + buffer` interior layers of a
    box: the fields of those layers and the ghost layers below are the first
    elements of every (x-fastest) field array, so the slab reads the box's
    own arrays.  The `buffer` layers above the kept ones carry an optical
    depth large enough that the approximate top boundary (I = S) does not
    reach the kept layers; apply() overwrites the kept layers of a Q_rad
    field with the slab's."""

    def __init__(self, grid, layers: int, buffer: int, num_rays: int = 24):
        ls = layers + buffer
        z0, z1 = grid.bounds[2]
        dz = (z1 - z0) / grid.k_size
        self.grid = Grid(grid.i_size, grid.j_size, ls, grid.k_margin,
                         bounds=(grid.bounds[0], grid.bounds[1], (z0, z0 + ls * dz)),
                         dtype=grid.dtype, device=grid.device)
        assert (self.grid.i_total, self.grid.j_total) == (grid.i_total, grid.j_total)
        self.layers, self.buffer = int(layers), int(buffer)
        self.rt = RT(self.grid, num_rays)
        self.iters = []

    def warm_start(self, full_rt) -> None:
        """Take the slab's intensities from a solve of the whole box: every
        ray's field keeps k slowest in both of its layouts, so the slab's
        planes are the first ones of the box's fields."""
        n = self.rt.rint[0].numel()
        self.rt.rint.view(self.rt.num_rays, -1).copy_(
            full_rt.rint.view(full_rt.num_rays, -1)[:, :n])

    def apply(self, ctx, qq, gm, kappa0, s_rad, q_full, **rt_kw):
        self.rt.source_from_fields(ctx, qq, gm, kappa0, s_rad)
        it, _ = self.rt.solve(ctx, **dict(rt_kw, **SLAB_FACES))
        self.iters.append(it)
        q = self.rt.qrad(ctx, **dict(rt_kw, **SLAB_FACES))
        m = self.grid.k_margin
        q_full[m:m + self.layers].copy_(q[m:m + self.layers])


def slab_layers(resid2, kdz, share: float, tau_buffer: float, max_layers: int):
    """(kept, buffer) layers of the bottom slab from the network's residual
    at a solve: kept = the fewest bottom layers holding `share` of the
    residual's squared norm, buffer = the fewest layers above them whose
    optical depth (sum of mean kappa dz) reaches `tau_buffer`; (0, 0) when
    that does not fit in `max_layers` (the residual is not at the bottom)."""
    r = np.asarray(resid2, dtype=np.float64)
    tot = r.sum()
    if tot <= 0.0:
        return 0, 0
    kept = int(np.searchsorted(np.cumsum(r) / tot, share) + 1)
    tau, buf = 0.0, 0
    while tau < tau_buffer and kept + buf < len(kdz):
        tau += float(kdz[kept + buf])
        buf += 1
    if tau < tau_buffer or kept + buf > max_layers:
        return 0, 0
    return kept, buf


@dataclass
class Stratification:
    t_bot: float = 2.0
    t_top: float = 0.5
    h_frac: float = 0.1
    amp: float = 1e-2
    gm: float = 5.0 / 3.0
    lz: float = 1.0
    tau_mid: float = 1.0
    t_relax: float = 0.05   # shortest radiative relaxation time, code units

    @property
    def grav(self) -> float:
        """g so that the pressure scale height T/g at the mean temperature is
        h_frac of the box height."""
        return 0.5 * (self.t_bot + self.t_top) / (self.h_frac * self.lz)

    def temperature(self, z):
        return self.t_bot + (self.t_top - self.t_bot) * np.asarray(z) / self.lz

    def pressure(self, z):
        """Hydrostatic dp/dz = -g p / T with T linear in z, p(0) = t_bot:
        p = t_bot (T / t_bot)^(g lz / (t_bot - t_top))."""
        ex = self.grav * self.lz / (self.t_bot - self.t_top)
        return self.t_bot * (self.temperature(z) / self.t_bot) ** ex

    def density(self, z):
        return self.pressure(z) / self.temperature(z)

    def kappa0(self, z_cells, dz) -> float:
        """kappa0 with sum over interior cells above mid-height of
        kappa0 ro T^3 dz equal to tau_mid."""
        z = np.asarray(z_cells)
        up = z >= 0.5 * self.lz
        return self.tau_mid / float(np.sum(self.density(z[up]) * self.temperature(z[up]) ** 3 * dz))

    def s_rad(self, z_cells, kappa0: float) -> float:
        """S = s_rad T^4 with the cooling time e ro / (4 pi kappa S) no
        shorter than t_relax anywhere in the box."""
        z = np.asarray(z_cells)
        T, ro = self.temperature(z), self.density(z)
        e = T / (self.gm - 1.0)
        tc = e * ro / (4.0 * math.pi * kappa0 * ro * T ** 3 * T ** 4)
        return float(np.min(tc) / self.t_relax)


def stratified_box(grid, strat: Stratification, seed: int, device="cpu"):
    """Initial condition (9, It, Jt, Kt), float64, MISO index order, ghosts
    included: the hydrostatic profile at every cell centre, plus seeded
    random-phase velocity (modes 1 <= |k| <= 4) and solenoidal magnetic
    perturbations of rms `amp` (x and y periodic, the z pattern spans the box
    height); the z ghost cells carry no velocity."""
    import torch
    z = np.asarray(grid.z, dtype=np.float64)
    shape = grid.shape
    q = torch.zeros((9,) + shape, dtype=torch.float64, device=device)
    T = torch.as_tensor(strat.temperature(z), device=device)[None, None, :]
    ro = torch.as_tensor(strat.density(z), device=device)[None, None, :]
    q[RO] = ro.expand(shape)
    q[EI] = (T / (strat.gm - 1.0)).expand(shape)
    if strat.amp > 0:
        rng = np.random.default_rng(seed)
        n = (grid.i_size, grid.j_size, grid.k_size)
        for comp0, sol in ((VX, False), (BX, True)):
            f = random_modes_field(n, 1.0, 4.0, rng, sol, device)
            rms = torch.sqrt(sum((f[c] ** 2).mean() for c in range(3)))
            for c in range(3):
                q[comp0 + c] = _periodic_pad(f[c] * (strat.amp / rms), grid)
        mk = grid.k_margin
        if mk:
            q[VX:VZ + 1, :, :, :mk] = 0.0
            q[VX:VZ + 1, :, :, -mk:] = 0.0
    return q


class RadiativeMHD:
    """Stratified radiative MHD box on one GPU."""

    def __init__(self, grid, ctx, strat: Stratification | None = None, cfl_number: float = 0.4,
                 rt_tolerance: float = 1e-6, rt_max_iters: int = 200):
        self.grid, self.ctx = grid, ctx
        self.strat = strat or Stratification()
        s = self.strat
        zc = np.asarray(grid.z)[grid.k_margin:grid.k_margin + grid.k_size]
        dz = float(grid.dz[0])
        self.kappa0 = s.kappa0(zc, dz)
        self.s_rad = s.s_rad(zc, self.kappa0)
        self.config = Config({"mhd": {"cfl_number": cfl_number}})
        self.model = MHD(self.config, grid, ctx, s.gm, src_kind=1, grav_axis=2, grav=s.grav)
        self.reservoir = None
        self.bc = None
        self.rt = RT(grid, 24)
        self.rt_kw = dict(RT_FACES, tolerance=rt_tolerance, max_iters=rt_max_iters)
        self.qrad = self.rt.qrad_field       # the Q_rad the next heating uses
        self.rt_iters = []
        self.surrogate = None
        self.flags = None
        self.q_rt = None                     # last solved Q_rad (learned mode)
        self.slab = None                     # bottom slab solved every step
        self.steps_done = 0
        # heating and the next CFL inside the step's last AV sweep (same bits
        # as the separate heating, ghost-fill and CFL passes)
        self.fused_heating = True
        self._next_dt = None

    # -- set-up -------------------------------------------------------------
    def initialize(self, seed: int) -> None:
        q = stratified_box(self.grid, self.strat, seed, device=self.grid.device)
        self.model.qq.load_ijk(q)
        # the z ghost cells keep the initial hydrostatic profile at rest
        if self.reservoir is None:
            from .mhd import Fields
            self.reservoir = Fields(self.grid)
        self.reservoir.data.copy_(self.model.qq.data)
        self.bc = BoundaryCondition(periodic=(True, True, "fixed"), fixed=self.reservoir)
        self.model.integrator.apply_boundary_condition(self.bc, self.model.qq)
        self._next_dt = None

    # -- radiative transfer -------------------------------------------------
    def solve_rt(self, **kw):
        """Plain RT: opacity and source from the fields, a (warm-started)
        solve (the run's RT settings, overridden by kw), Q_rad into
        self.rt.qrad_field."""
        self.rt.source_from_fields(self.ctx, self.model.qq, self.strat.gm, self.kappa0, self.s_rad)
        it, md = self.rt.solve(self.ctx, **dict(self.rt_kw, **kw))
        self.rt_iters.append(it)
        self.rt.qrad(self.ctx, **self.rt_kw)
        return self.rt.qrad_field

    def reference_qrad(self):
        """Q_rad of the current state from a separate, fully converged RT
        solve (its own intensities, so the run's warm start is untouched)."""
        if getattr(self, "_rt_ref", None) is None:
            self._rt_ref = RT(self.grid, 24)
        r = self._rt_ref
        r.source_from_fields(self.ctx, self.model.qq, self.strat.gm, self.kappa0, self.s_rad)
        r.solve(self.ctx, **self.rt_kw)
        return r.qrad(self.ctx, **self.rt_kw)

    def attach_surrogate(self, surrogate, rt_every: int = 16, flag_tol: float | None = None,
                         correct: bool = True, slab_share: float | None = 0.99,
                         slab_tau: float = 10.0, slab_max_frac: float = 0.25,
                         slab_tolerance: float | None = None,
                         slab_max_iters: int | None = None,
                         solve_max_iters: int | None = None) -> None:
        """Learned RT: the network every step, a plain solve every rt_every
        steps.  correct: add the column correction the last solve left (the
        solve minus the network on the same state), so the network carries
        the change since the solve; flag_tol: columns whose network misses
        the solve by more than this (relative L2) take the solved values
        until the next solve (None: no flags); slab_share: at every solve
        the bottom layers holding this share of the network's squared
        residual, topped by slab_tau of optical depth, are handed to a
        plain slab solve on every step (to slab_tolerance, default the
        run's RT tolerance, in at most slab_max_iters iterations, default
        the run's limit), if they fit in slab_max_frac of the height (None:
        no slab); solve_max_iters caps the iterations of the plain solve
        every rt_every steps (default: the run's RT limit)."""
        import torch
        self.surrogate = surrogate
        self.rt_every = int(rt_every)
        self.flag_tol = flag_tol
        self.correct = bool(correct)
        self.slab_share, self.slab_tau, self.slab_max_frac = slab_share, slab_tau, slab_max_frac
        self.slab_kw = dict(self.rt_kw)
        if slab_tolerance is not None:
            self.slab_kw["tolerance"] = float(slab_tolerance)
        if slab_max_iters is not None:
            self.slab_kw["max_iters"] = int(slab_max_iters)
        self.solve_kw = {} if solve_max_iters is None else {"max_iters": int(solve_max_iters)}
        self.slab = None
        ncol = self.grid.i_size * self.grid.j_size
        self.flags = torch.zeros(ncol, dtype=torch.int32, device=self.grid.device)
        self.q_rt = torch.zeros_like(self.rt.qrad_field)
        self.q_mlp = torch.zeros_like(self.rt.qrad_field)
        self.corr = torch.zeros_like(self.rt.qrad_field) if self.correct else None

    def learned_qrad(self, out=None, correct=None, use_flags=None):
        """Network Q_rad of the current state (into self.q_mlp by default),
        with the run's correction and flags unless told otherwise."""
        correct = self.correct if correct is None else correct
        use_flags = (self.flag_tol is not None) if use_flags is None else use_flags
        out = self.q_mlp if out is None else out
        self.surrogate.predict(self.grid, self.model.qq, self.strat.gm, out,
                               flags=self.flags if use_flags else None, q_rt=self.q_rt,
                               corr=self.corr if correct else None)
        return out

    # -- stepping -----------------------------------------------------------
    def _advance(self, qrad) -> float:
        integ = self.model.integrator
        if self.fused_heating:
            dt = self._next_dt if self._next_dt is not None else integ.cfl(self.model.qq)
            self._next_dt = integ.update_heated(dt, self.bc, self.model.qq, qrad)
        else:
            dt = integ.step(self.bc, self.model.qq)
            self.rt.heat(self.ctx, self.model.qq, dt, qrad)
            integ.apply_boundary_condition(self.bc, self.model.qq)
        self.steps_done += 1
        return dt

    def step_plain(self) -> float:
        q = self.solve_rt()
        return self._advance(q)

    def step_learned(self, probe=None) -> float:
        """One step with the learned Q_rad (a plain solve every rt_every
        steps).  probe(q, solved) is called with the Q_rad about to be used,
        before the state advances."""
        if self.steps_done % self.rt_every == 0:
            q = self.solve_rt(**self.solve_kw)
            self.q_rt.copy_(q)
            # network on the same state -> correction and residual flags
            self.surrogate.predict(self.grid, self.model.qq, self.strat.gm, self.q_mlp,
                                   flags=None, q_rt=None)
            tol = self.flag_tol if self.flag_tol is not None else float("inf")
            self.surrogate.residual(self.grid, self.q_rt, self.flags, tol, self.corr)
            if self.slab_share is not None:
                self._choose_slab()
            if probe:
                probe(self.q_rt, True)
            return self._advance(self.q_rt)
        q = self.learned_qrad()
        if self.slab is not None:
            self.slab.apply(self.ctx, self.model.qq, self.strat.gm, self.kappa0, self.s_rad, q,
                            **self.slab_kw)
        if probe:
            probe(q, False)
        return self._advance(q)

    def _choose_slab(self) -> None:
        """Bottom slab from the network's residual on the just-solved state
        (self.q_mlp against self.q_rt, per height) and the layers' mean
        optical thickness; rebuilt only when the kept layers grow or halve."""
        g = self.grid
        m = g.k_margin
        inner = (slice(m, m + g.k_size), slice(g.j_margin, g.j_margin + g.j_size),
                 slice(g.i_margin, g.i_margin + g.i_size))
        d = (self.q_mlp - self.q_rt)[inner]
        resid2 = (d * d).sum(dim=(1, 2)).cpu().numpy()
        kdz = (self.rt.abs_coeff[inner].mean(dim=(1, 2)) * float(g.dz[m])).cpu().numpy()
        kept, buf = slab_layers(resid2, kdz, self.slab_share, self.slab_tau,
                                int(self.slab_max_frac * g.k_size))
        cur = self.slab
        if kept == 0:
            self.slab = None
        elif cur is None or kept > cur.layers or 2 * kept < cur.layers or buf != cur.buffer:
            self.slab = BottomSlabRT(g, kept, buf)
        if self.slab is not None:
            # the solve just made is the best start for the slab's next solve
            self.slab.warm_start(self.rt)


def rel_l2(a, b, grid) -> float:
    """Relative L2 of a against b over the interior cells."""
    k0, j0, i0 = grid.k_margin, grid.j_margin, grid.i_margin
    sl = (slice(k0, k0 + grid.k_size), slice(j0, j0 + grid.j_size), slice(i0, i0 + grid.i_size))
    d = (a[sl] - b[sl]).double()
    return float(d.norm() / b[sl].double().norm())
