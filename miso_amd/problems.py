"""Initial conditions of the benchmark workloads (MISO's `InitialCondition::apply`
functors, written for this backend).

Every generator returns a float64 torch tensor of shape (9, i_total, j_total,
k_total) in MISO's (i, j, k) index order, ghosts included, on the requested
torch device; `Fields.copy_from_host` / `load_ijk` place it on the GPU.
Field order: ro vx vy vz bx by bz ei ph.
"""
from __future__ import annotations

import math

import numpy as np

RO, VX, VY, VZ, BX, BY, BZ, EI, PH = range(9)


def _torch():
    import torch
    return torch


def _coords(grid, device):
    torch = _torch()
    x = torch.as_tensor(np.asarray(grid.x, dtype=np.float64), device=device)
    y = torch.as_tensor(np.asarray(grid.y, dtype=np.float64), device=device)
    z = torch.as_tensor(np.asarray(grid.z, dtype=np.float64), device=device)
    return x[:, None, None], y[None, :, None], z[None, None, :]


def orszag_tang(grid, gm=5.0 / 3.0, device="cpu"):
    """2D Orszag-Tang vortex: rho = 25/(36 pi), p = 5/(12 pi),
    v = (-sin 2 pi y, sin 2 pi x), B = (-sin 2 pi y, sin 4 pi x)/sqrt(4 pi),
    evaluated at every cell centre including ghosts (the layout of
    demo/mhd2d_vortex/src/main.cpp:19-39 with the benchmark's constants)."""
    torch = _torch()
    x, y, z = _coords(grid, device)
    shape = grid.shape
    q = torch.zeros((9,) + shape, dtype=torch.float64, device=device)
    ro = 25.0 / (36.0 * math.pi)
    pr = 5.0 / (12.0 * math.pi)
    b0 = 1.0 / math.sqrt(4.0 * math.pi)
    q[RO] = ro
    q[EI] = pr / (gm - 1.0) / ro
    q[VX] = (-torch.sin(2 * math.pi * y)).expand(shape)
    q[VY] = torch.sin(2 * math.pi * x).expand(shape)
    q[BX] = (-b0 * torch.sin(2 * math.pi * y)).expand(shape)
    q[BY] = (b0 * torch.sin(4 * math.pi * x)).expand(shape)
    return q


def _half_space_modes(kmin, kmax):
    modes = []
    r = int(math.ceil(kmax))
    for kx in range(-r, r + 1):
        for ky in range(-r, r + 1):
            for kz in range(-r, r + 1):
                k2 = kx * kx + ky * ky + kz * kz
                if not (kmin * kmin <= k2 <= kmax * kmax):
                    continue
                if kz > 0 or (kz == 0 and (ky > 0 or (ky == 0 and kx > 0))):
                    modes.append((kx, ky, kz))
    return np.array(modes, dtype=np.int64)


def _periodic_pad(f, grid):
    """Interior field (ni, nj, nk) -> (It, Jt, Kt) with periodic ghosts."""
    torch = _torch()
    idx = []
    for n, m in ((grid.i_size, grid.i_margin), (grid.j_size, grid.j_margin),
                 (grid.k_size, grid.k_margin)):
        idx.append(torch.as_tensor((np.arange(n + 2 * m) - m) % n, device=f.device))
    return f[idx[0]][:, idx[1]][:, :, idx[2]]


def random_modes_field(n, kmin, kmax, rng, solenoidal, device):
    """Three components of a random-phase field made of the integer modes
    kmin <= |k| <= kmax on a periodic n[0] x n[1] x n[2] box (flat amplitude,
    uniform phases); optionally projected onto k-perpendicular (div-free)."""
    torch = _torch()
    modes = _half_space_modes(kmin, kmax)
    # complex amplitude vectors: random direction components with random phases
    amp = rng.normal(size=(len(modes), 3)) + 1j * rng.normal(size=(len(modes), 3))
    if solenoidal:
        k = modes.astype(np.float64)
        kk = (k * k).sum(1, keepdims=True)
        amp = amp - k * (amp * k).sum(1, keepdims=True) / kk
    spec = torch.zeros((3,) + tuple(n), dtype=torch.complex128, device=device)
    ii = torch.as_tensor(modes % np.array(n), device=device)
    jj = torch.as_tensor((-modes) % np.array(n), device=device)
    a = torch.as_tensor(amp.T.copy(), device=device)
    for c in range(3):
        spec[c, ii[:, 0], ii[:, 1], ii[:, 2]] = a[c]
        spec[c, jj[:, 0], jj[:, 1], jj[:, 2]] = a[c].conj()
    f = torch.fft.ifftn(spec, dim=(1, 2, 3)).real
    return f


def turbulence(grid, gm=5.0 / 3.0, seed=2024, v_rms=0.5, b_rms=0.3, kmin=1, kmax=4,
               device="cpu"):
    """3D decaying MHD turbulence: rho = 1, p = 1, v and B seeded random-phase
    modes 1 <= |k| <= 4, B divergence-free by projection, rms-normalised to
    v_rms / b_rms over the interior; periodic box."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    n = (grid.i_size, grid.j_size, grid.k_size)
    v = random_modes_field(n, kmin, kmax, rng, False, device)
    b = random_modes_field(n, kmin, kmax, rng, True, device)
    v = v * (v_rms / torch.sqrt((v * v).sum(0).mean()))
    b = b * (b_rms / torch.sqrt((b * b).sum(0).mean()))
    q = torch.zeros((9,) + grid.shape, dtype=torch.float64, device=device)
    q[RO] = 1.0
    q[EI] = 1.0 / (gm - 1.0) / 1.0
    for c in range(3):
        q[VX + c] = _periodic_pad(v[c], grid)
        q[BX + c] = _periodic_pad(b[c], grid)
    return q


def kelvin_helmholtz(grid, gm=1.4, seed=0, pr=2.5, ro_upper=1.0, ro_lower=2.0, dv=1.0,
                     width=0.02, bx=None, v_amp=1e-3, device="cpu"):
    """Kelvin-Helmholtz shear layer on [-0.5, 0.5]^2: tanh profile of width
    `width` in y, density ratio ro_lower/ro_upper, velocity jump dv, uniform
    Bx (seeded in [0, 0.2] when not given), 1e-3 seeded random vy
    (pressure and gamma as demo/hd2d_kelvin_helmholtz/config.yaml)."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    if bx is None:
        bx = float(rng.uniform(0.0, 0.2))
    x, y, z = _coords(grid, device)
    shape = grid.shape
    q = torch.zeros((9,) + shape, dtype=torch.float64, device=device)
    t = torch.tanh(y / width).expand(shape)
    q[RO] = 0.5 * (ro_upper + ro_lower) + 0.5 * (ro_upper - ro_lower) * t
    q[VX] = 0.5 * dv * t
    noise = torch.as_tensor(rng.uniform(-1.0, 1.0, size=shape), device=device)
    q[VY] = v_amp * noise
    q[BX] = bx
    q[EI] = pr / (gm - 1.0) / q[RO]
    return q


def shock_tube(grid, axis=0, gm=1.4, rol=1.0, prl=1.0, vvl=0.0, ror=0.125, prr=0.1, vvr=0.0,
               device="cpu"):
    """Sod shock tube along `axis` (the reference's only end-to-end test,
    miso/tests/serial/test_hd1d_shock_tube/initial_condition.hpp:147-184:
    the state jumps at 0.5 and the velocity goes to vx)."""
    torch = _torch()
    x, y, z = _coords(grid, device)
    s = (x, y, z)[axis].expand(grid.shape)
    q = torch.zeros((9,) + grid.shape, dtype=torch.float64, device=device)
    left = s < 0.5
    q[RO] = torch.where(left, torch.tensor(rol, dtype=torch.float64, device=device),
                        torch.tensor(ror, dtype=torch.float64, device=device))
    q[EI] = torch.where(left, torch.tensor(prl / (gm - 1.0) / rol, dtype=torch.float64,
                                           device=device),
                        torch.tensor(prr / (gm - 1.0) / ror, dtype=torch.float64, device=device))
    q[VX] = torch.where(left, torch.tensor(vvl, dtype=torch.float64, device=device),
                        torch.tensor(vvr, dtype=torch.float64, device=device))
    return q


def to_numpy(q, dtype=np.float64) -> np.ndarray:
    """(9, It, Jt, Kt) MISO-order numpy array of the run's Real type."""
    return np.ascontiguousarray(q.detach().cpu().numpy().astype(dtype))
