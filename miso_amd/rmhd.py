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
                   residual flags (columns where the network misses the
                   solve by more than `flag_tol` in relative L2) that take the
                   last solved Q_rad until the next solve.

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

from .mhd import BoundaryCondition, Config, MHD
from .problems import random_modes_field, _periodic_pad
from .rt import RT

RO, VX, VY, VZ, BX, BY, BZ, EI, PH = range(9)

RT_FACES = dict(periodic=(True, True, False), face_kind=(0, 0, 0, 0, 1, 2))


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
        self.steps_done = 0

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

    # -- radiative transfer -------------------------------------------------
    def solve_rt(self):
        """Plain RT: opacity and source from the fields, a (warm-started)
        solve, Q_rad into self.rt.qrad_field."""
        self.rt.source_from_fields(self.ctx, self.model.qq, self.strat.gm, self.kappa0, self.s_rad)
        it, md = self.rt.solve(self.ctx, **self.rt_kw)
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
                         correct: bool = True) -> None:
        """Learned RT: the network every step, a plain solve every rt_every
        steps.  correct: add the column correction the last solve left (the
        solve minus the network on the same state), so the network carries
        the change since the solve; flag_tol: columns whose network misses
        the solve by more than this (relative L2) take the solved values
        until the next solve (None: no flags)."""
        import torch
        self.surrogate = surrogate
        self.rt_every = int(rt_every)
        self.flag_tol = flag_tol
        self.correct = bool(correct)
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
            q = self.solve_rt()
            self.q_rt.copy_(q)
            # network on the same state -> correction and residual flags
            self.surrogate.predict(self.grid, self.model.qq, self.strat.gm, self.q_mlp,
                                   flags=None, q_rt=None)
            tol = self.flag_tol if self.flag_tol is not None else float("inf")
            self.surrogate.residual(self.grid, self.q_rt, self.flags, tol, self.corr)
            if probe:
                probe(self.q_rt, True)
            return self._advance(self.q_rt)
        q = self.learned_qrad()
        if probe:
            probe(q, False)
        return self._advance(q)


def rel_l2(a, b, grid) -> float:
    """Relative L2 of a against b over the interior cells."""
    k0, j0, i0 = grid.k_margin, grid.j_margin, grid.i_margin
    sl = (slice(k0, k0 + grid.k_size), slice(j0, j0 + grid.j_size), slice(i0, i0 + grid.i_size))
    d = (a[sl] - b[sl]).double()
    return float(d.norm() / b[sl].double().norm())
