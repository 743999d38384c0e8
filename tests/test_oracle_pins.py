"""Pins of the CPU oracle (oracle/mhd_oracle.c) -- what makes it trustworthy.

MISO ships no golden vectors for the MHD step, so the oracle is pinned by
(a) the reference's own headers compiled in oracle/_ref (limiter, float-
narrowing min/max, constants), (b) the reference's only end-to-end test
(Sod, miso/tests/serial/test_hd1d_shock_tube.cpp:52-81) and its boundary
KATs (test_boundary_condition_common.hpp:15-120), and (c) what the
mathematics fixes: the 4th-order stencil is exact on quartics, linear
sound / Alfven / fast / divergence-cleaning waves of the semi-discrete system
grow by the RK4 polynomial of their exact discrete eigenvalues, and the flux
form conserves mass, momentum, field and total energy.  test_planted_mistakes
builds the oracle with deliberate errors and shows each is caught.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")


# ------------------------------------------------------------------ helpers
def lib_ref():
    if not os.path.exists(O.REF_PATH):
        if O.build_ref() is None:
            pytest.skip("reference headers not present and oracle/_ref not built")
    L = C.CDLL(O.REF_PATH)
    for S, R in (("f64", C.c_double), ("f32", C.c_float)):
        getattr(L, f"ref_dqq_eval_{S}").argtypes = [R] * 4
        getattr(L, f"ref_dqq_eval_{S}").restype = R
        getattr(L, f"ref_flux_core_{S}").argtypes = [R] * 6
        getattr(L, f"ref_flux_core_{S}").restype = R
        getattr(L, f"ref_const_{S}").argtypes = [C.c_int]
        getattr(L, f"ref_const_{S}").restype = R
    for n, k in (("ref_min2_f64", 2), ("ref_max3_f64", 3), ("ref_min3_f64", 3)):
        getattr(L, n).argtypes = [C.c_double] * k
        getattr(L, n).restype = C.c_double
    return L


class OracleVariant:
    """The oracle module's functions bound to another build of the library."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        self._saved = O._lib
        O._lib = None
        saved_path = O.LIB_PATH
        O.LIB_PATH = self.path
        try:
            O.lib()
        finally:
            O.LIB_PATH = saved_path
        return O

    def __exit__(self, *a):
        O._lib = self._saved


def planted(n):
    out = os.path.join(ORACLE_DIR, "_build", f"libmiso_oracle_plant{n}.so")
    srcs = [os.path.join(ORACLE_DIR, f) for f in ("mhd_oracle.c", "rt_oracle.c")]
    subprocess.run(["gcc", "-std=gnu11", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC",
                    "-shared", f"-DOG_PLANT={n}", "-o", out] + srcs + ["-lm"], check=True)
    return OracleVariant(out)


def uniform(grid, ro=1.0, v=(0, 0, 0), b=(0, 0, 0), ei=1.5, dtype=np.float64):
    q = np.zeros((9,) + grid.shape, dtype=dtype)
    q[0] = ro
    for c in range(3):
        q[1 + c] = v[c]
        q[4 + c] = b[c]
    q[7] = ei
    return q


def k_eff(k, h):
    """Symbol of the 4th-order central difference on exp(ikx)."""
    return (8 * np.sin(k * h) - np.sin(2 * k * h)) / (6 * h)


def rk4_poly(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def mode_coeff(f, axis, m, n):
    """Fourier coefficient of mode m along `axis` (interior of length n)."""
    x = np.arange(n)
    e = np.exp(-2j * np.pi * m * x / n)
    shape = [1, 1, 1]
    shape[axis] = n
    return (f * e.reshape(shape)).sum() / f.size


NOAV = dict(cs_fac=0.0, ca_fac=0.0, vv_fac=0.0)


def _interior(grid):
    g = grid.g
    return (slice(g.i_margin, g.i_total - g.i_margin), slice(g.j_margin, g.j_total - g.j_margin),
            slice(g.k_margin, g.k_total - g.k_margin))


# ------------------------------------------------------- (a) reference headers
def test_limiter_matches_reference_headers():
    L = lib_ref()
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.standard_normal(4000), rng.standard_normal(2000) * 1e-6,
                           [0.0, -0.0, 1e-30, -1e-30, 1.0, 1.0 + 1e-12]])
    for S, dt in (("f64", np.float64), ("f32", np.float32)):
        for t in range(3000):
            a, b, c = (dt(x) for x in rng.choice(vals, 3))
            ep = dt(rng.choice([1.0, 1.5, 2.0]))
            r = getattr(L, f"ref_dqq_eval_{S}")(a, b, c, ep)
            o = O.dqq_eval(a, b, c, ep, dtype=dt)
            assert r == o or (math.isnan(r) and math.isnan(o)), (S, a, b, c, ep, r, o)
            qd, qu, dd, du = (dt(x) for x in rng.choice(vals, 4))
            cc, fh = dt(abs(rng.standard_normal()) + 0.1), dt(rng.choice([0.0, 0.5, 1.0]))
            r = getattr(L, f"ref_flux_core_{S}")(qd, qu, dd, du, cc, fh)
            o = O.flux_core(qd, qu, dd, du, cc, fh, dtype=dt)
            assert r == o, (S, qd, qu, dd, du, cc, fh, r, o)


def test_float_narrowing_and_constants_match_reference():
    L = lib_ref()
    # util::min2<double> rounds through float (utility.hpp:91-124)
    assert L.ref_min2_f64(1.0 + 1e-12, 2.0) == 1.0
    assert L.ref_min3_f64(0.1, 0.2, 0.3) == float(np.float32(0.1))
    for w, v in enumerate((1.0 / 12.0, 1.0 / (4 * math.pi), 1.0 / (8 * math.pi))):
        assert L.ref_const_f64(w) == v
        assert L.ref_const_f32(w) == float(np.float32(v))


def test_oracle_matches_committed_reference_vectors():
    """The limiter, constants and quadrature of the oracle against vectors
    the reference's own headers produced (tests/golden/make_ref_vectors.py;
    this check needs neither the reference nor oracle/_ref)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "ref_limiter_quadrature.npz")
    d = np.load(path)
    for S, dt in (("f64", np.float64), ("f32", np.float32)):
        for row, r in zip(d[f"dqq_in_{S}"], d[f"dqq_out_{S}"]):
            o = O.dqq_eval(*row, dtype=dt)
            assert dt(o) == r or (np.isnan(o) and np.isnan(r)), (S, row, o, r)
        for row, r in zip(d[f"flux_in_{S}"], d[f"flux_out_{S}"]):
            assert dt(O.flux_core(*row, dtype=dt)) == r, (S, row)
    # the narrowing min/max the kernels reproduce: (double)fminf((float)a, (float)b)
    f = lambda *xs: [float(np.float32(x)) for x in xs]
    for (a, b, c), m2, m3, x3 in zip(d["minmax_in"], d["min2_out"], d["min3_out"], d["max3_out"]):
        assert min(f(a, b)) == m2 and min(f(a, b, c)) == m3 and max(f(a, b, c)) == x3
    w, mx, my, mz = O.rt_quadrature(24)
    q = d["quadrature_24"]
    assert np.array_equal(q, np.stack([w, mx, my, mz], 1))
    inv12, pii4, pii8 = d["const_f64"][:3]
    assert (inv12, pii4, pii8) == (1.0 / 12.0, 1.0 / (4 * math.pi), 1.0 / (8 * math.pi))


def minmod_slope(a, b, c, ep):
    """minmod(ep*(b-a), (c-a)/2, ep*(c-b)) on float-rounded arguments."""
    args = [np.float32(ep * (b - a)), np.float32(ep * (c - b)), np.float32((c - a) * 0.5)]
    if all(x > 0 for x in args):
        return float(min(args))
    if all(x < 0 for x in args):
        return float(max(args))
    return 0.0


def test_limiter_is_generalised_minmod():
    """dqq_eval = minmod(ep*dq_-, (q_+ - q_-)/2, ep*dq_+) on float-rounded
    arguments (Rempel 2014 ApJ 789 132 eq. 8 with epsilon = ep); flux_core
    with fh = 0 switches on the sign of the limited ratio only."""
    rng = np.random.default_rng(1)
    for _ in range(2000):
        a, b, c = rng.standard_normal(3)
        ep = rng.choice([1.0, 1.5, 2.0])
        assert O.dqq_eval(a, b, c, ep) == minmod_slope(a, b, c, ep)
    # smooth monotone data: reconstruction meets in the middle -> no flux
    assert O.flux_core(1.0, 2.0, 1.0, 1.0, 3.0, 1.0) == 0.0
    # extremum: slopes vanish, flux = -0.5*cc*(q_up - q_dw)
    assert O.flux_core(1.0, 2.0, 0.0, 0.0, 3.0, 1.0) == -0.5 * 3.0 * 1.0


# ------------------------------------------------------------ (b) reference KATs
def test_ghost_fill_reference_kat():
    g = O.make_grid((10, 11, 12), 2, bounds=((0, 1), (0, 2), (0, 3)))
    It, Jt, Kt = g.shape
    for sign in (1, -1):
        q = np.zeros((9,) + g.shape)
        q[0] = np.arange(It, dtype=np.float64)[:, None, None]
        s = [[1] * 9 for _ in range(3)]
        s[0][0] = sign
        O.ghost_fill(g, O.Params(periodic=(0, 1, 1), sym_sign=s), q)
        ro = q[0]
        jk = (slice(2, -2), slice(2, -2))
        assert np.array_equal(ro[(0,) + jk], sign * ro[(3,) + jk])
        assert np.array_equal(ro[(1,) + jk], sign * ro[(2,) + jk])
        assert np.array_equal(ro[(It - 1,) + jk], sign * ro[(It - 4,) + jk])
        assert np.array_equal(ro[(It - 2,) + jk], sign * ro[(It - 3,) + jk])
    # periodic: ghost = interior image (brute force)
    q = np.random.default_rng(2).standard_normal((9,) + g.shape)
    O.ghost_fill(g, O.Params(periodic=(1, 1, 1)), q)
    n = (10, 11, 12)
    for (i, j, k) in [(0, 5, 5), (1, 6, 7), (It - 1, 5, 5), (5, 0, 5), (5, Jt - 2, 6), (4, 4, 0),
                      (4, 4, Kt - 1)]:
        src = [(x - 2) % nn + 2 for x, nn in zip((i, j, k), n)]
        assert np.array_equal(q[:, i, j, k], q[:, src[0], src[1], src[2]])


def test_ghost_fill_fixed_axis_takes_the_fixed_values():
    """periodic[a] = 2: every fill sets the axis' ghost planes from the
    fixed fields; the x and y halo exchange then wraps the cells they share
    with those planes (the user functor runs before the exchange); a
    mirrored axis reflects the fixed values it meets."""
    g = O.make_grid((6, 5, 7), 2)
    It, Jt, Kt = g.shape
    rng = np.random.default_rng(3)
    q = rng.standard_normal((9,) + g.shape)
    fx = rng.standard_normal((9,) + g.shape)
    q0 = q.copy()
    O.ghost_fill(g, O.Params(periodic=(1, 1, 2), fixed=fx), q)
    ii, jj = slice(2, -2), slice(2, -2)
    for k in (0, 1, Kt - 2, Kt - 1):
        assert np.array_equal(q[:, ii, jj, k], fx[:, ii, jj, k])
        # x-ghost / z-ghost edge: periodic image along x of the fixed plane
        assert np.array_equal(q[:, 0, 3, k], fx[:, It - 4, 3, k])
        assert np.array_equal(q[:, It - 1, 3, k], fx[:, 3, 3, k])
    # interior untouched, x faces wrap as with a periodic z
    assert np.array_equal(q[:, ii, jj, 2:-2], q0[:, ii, jj, 2:-2])
    assert np.array_equal(q[:, 1, 3, 4], q0[:, It - 3, 3, 4])
    # a second fill after the ghosts were disturbed restores them
    q[:, ii, jj, 0] = 7.0
    O.ghost_fill(g, O.Params(periodic=(1, 1, 2), fixed=fx), q)
    assert np.array_equal(q[:, ii, jj, 0], fx[:, ii, jj, 0])
    # y mirror + z fixed: the (y ghost, z ghost) edge mirrors the fixed plane
    s = [[1] * 9 for _ in range(3)]
    s[1][2] = -1
    q = q0.copy()
    O.ghost_fill(g, O.Params(periodic=(1, 0, 2), sym_sign=s, fixed=fx), q)
    assert np.array_equal(q[2, 3, 0, 0], -fx[2, 3, 3, 0])
    assert np.array_equal(q[0, 3, 1, Kt - 1], fx[0, 3, 2, Kt - 1])


def sod_exact(x, t, x0, gm, rl, pl, ul, rr, pr, ur):
    """Exact Riemann solution of the Euler equations (Toro, Riemann Solvers
    and Numerical Methods for Fluid Dynamics, ch. 4)."""
    from scipy.optimize import brentq
    cl, cr = math.sqrt(gm * pl / rl), math.sqrt(gm * pr / rr)

    def f(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((gm + 1) * rk), (gm - 1) / (gm + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2 * ck / (gm - 1) * ((p / pk) ** ((gm - 1) / (2 * gm)) - 1)

    ps = brentq(lambda p: f(p, rl, pl, cl) + f(p, rr, pr, cr) + ur - ul, 1e-10, 10 * max(pl, pr))
    us = 0.5 * (ul + ur) + 0.5 * (f(ps, rr, pr, cr) - f(ps, rl, pl, cl))
    ro = np.empty_like(x)
    u = np.empty_like(x)
    for n, xi in enumerate(x):
        s = (xi - x0) / t
        if s < us:  # left of contact
            if ps > pl:
                sl = ul - cl * math.sqrt((gm + 1) / (2 * gm) * ps / pl + (gm - 1) / (2 * gm))
                ro[n], u[n] = ((rl * ((ps / pl + (gm - 1) / (gm + 1)) /
                                      ((gm - 1) / (gm + 1) * ps / pl + 1)), us) if s > sl
                               else (rl, ul))
            else:
                csl = cl * (ps / pl) ** ((gm - 1) / (2 * gm))
                if s < ul - cl:
                    ro[n], u[n] = rl, ul
                elif s > us - csl:
                    ro[n], u[n] = rl * (ps / pl) ** (1 / gm), us
                else:
                    c = 2 / (gm + 1) * (cl + (gm - 1) / 2 * (ul - s))
                    ro[n], u[n] = rl * (c / cl) ** (2 / (gm - 1)), 2 / (gm + 1) * (cl + (gm - 1) / 2 * ul + s)
        else:
            if ps > pr:
                sr = ur + cr * math.sqrt((gm + 1) / (2 * gm) * ps / pr + (gm - 1) / (2 * gm))
                ro[n], u[n] = ((rr * ((ps / pr + (gm - 1) / (gm + 1)) /
                                      ((gm - 1) / (gm + 1) * ps / pr + 1)), us) if s < sr
                               else (rr, ur))
            else:
                csr = cr * (ps / pr) ** ((gm - 1) / (2 * gm))
                if s > ur + cr:
                    ro[n], u[n] = rr, ur
                elif s < us + csr:
                    ro[n], u[n] = rr * (ps / pr) ** (1 / gm), us
                else:
                    c = 2 / (gm + 1) * (cr - (gm - 1) / 2 * (ur - s))
                    ro[n], u[n] = rr * (c / cr) ** (2 / (gm - 1)), 2 / (gm + 1) * (-cr + (gm - 1) / 2 * ur + s)
    return ro, u


def run_sod(axis, dtype=np.float32, n=128, tend=0.2):
    """The reference's Sod test set-up (test_hd1d_shock_tube/config/config_x.yaml:
    128 cells, gamma 1.4, default CFL and AV, mirror boundaries, t_end 0.2)."""
    sizes = [1, 1, 1]
    sizes[axis] = n
    g = O.make_grid(tuple(sizes), 2, dtype=dtype)
    x = g.axis("xyz"[axis])
    q = np.zeros((9,) + g.shape, dtype=dtype)
    shape = [1, 1, 1]
    shape[axis] = -1
    left = (x < 0.5).reshape(shape)
    gm = 1.4
    q[0] = np.where(left, 1.0, 0.125)
    q[7] = np.where(left, 1.0 / (gm - 1) / 1.0, 0.1 / (gm - 1) / 0.125)
    p = O.Params(gm=gm, periodic=(0, 0, 0))
    O.ghost_fill(g, p, q)
    time = dtype(0)
    while time < dtype(tend):
        dt = O.run(g, p, q, 1)[0]
        time = dtype(time + dtype(dt))
    return g, q, float(time)


def test_sod_shock_tube_kat():
    g, q, t = run_sod(0)
    gm = 1.4
    x = g.axis("x").astype(np.float64)[2:-2]
    ro_ex, u_ex = sod_exact(x, t, 0.5, gm, 1.0, 1.0, 0.0, 0.125, 0.1, 0.0)
    ro = q[0, 2:-2, 0, 0].astype(np.float64)
    vx = q[1, 2:-2, 0, 0].astype(np.float64)
    assert np.abs(ro - ro_ex).sum() / ro.sum() < 5e-2
    assert np.abs(vx - u_ex).sum() / vx.sum() < 5e-2
    # x/y/z runs agree as doctest::Approx(.).epsilon(1e-5).scale(1.0) asks
    # (test_hd1d_shock_tube.cpp:73-81)
    for axis in (1, 2):
        g2, q2, t2 = run_sod(axis)
        sl = [0, 0, 0]
        sl[axis] = slice(2, -2)
        a = q2[0][tuple(sl)].astype(np.float64)
        b = q[0, 2:-2, 0, 0].astype(np.float64)
        assert np.all(np.abs(a - b) < 1e-5 * (1.0 + np.maximum(np.abs(a), np.abs(b))))


# -------------------------------------------------------- (c) mathematics
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_stencil_exact_on_quartics(axis):
    """Continuity with uniform velocity: rho' = rho - dt*u*d(rho)/dx exactly
    for a quartic rho (the 5-point stencil is exact to degree 4); a quintic
    shows the stencil's known truncation error -h^4/30 f^(5)."""
    sizes = [1, 1, 1]
    sizes[axis] = 16
    g = O.make_grid(tuple(sizes), 2)
    x = g.axis("xyz"[axis])
    h = float(g.axis("d" + "xyz"[axis])[0])
    shape = [1, 1, 1]
    shape[axis] = -1
    u = 0.7
    dt = 1e-3
    for deg, coef in ((4, (0.3, -0.2, 0.5, 0.1, 0.25)), (5, (0, 0, 0, 0, 0, 0.4))):
        ro = sum(c * x ** m for m, c in enumerate(coef)) + 2.0
        dro = sum(m * c * x ** (m - 1) for m, c in enumerate(coef) if m)
        q = uniform(g, v=tuple(u if a == axis else 0 for a in range(3)))
        q[0] = ro.reshape(shape)
        rslt = np.zeros_like(q)
        O.rk_stage(g, O.Params(), dt, 0.0, 1.0, q, q, rslt)
        interior = [0, 0, 0]
        interior[axis] = slice(2, -2)
        got = (q[0] - rslt[0])[tuple(interior)] / (dt * u)
        want = dro[2:-2]
        if deg == 4:
            assert np.allclose(got, want, rtol=0, atol=1e-11), (got - want)
        else:
            trunc = -h ** 4 / 30 * 120 * coef[5]
            assert np.allclose(got - want, trunc, rtol=1e-6, atol=1e-13)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_uniform_state_is_steady(dtype):
    g = O.make_grid((12, 10, 8), 2, dtype=dtype)
    q = uniform(g, ro=1.3, v=(0.2, -0.1, 0.05), b=(0.1, 0.3, -0.2), ei=2.0, dtype=dtype)
    q0 = q.copy()
    O.update(g, O.Params(), 1e-3, q)
    tol = 1e-15 if dtype == np.float64 else 1e-6
    inter = (slice(None),) + _interior(g)
    assert np.max(np.abs(q[inter] - q0[inter])) <= tol * 3, np.max(np.abs(q[inter] - q0[inter]))


def _wave_run(axis, n, mode, init, lam_fn, steps=6, cfl=0.4, av=NOAV, mod=O):
    """Run `steps` updates of a small-amplitude eigenmode along `axis` and
    return (measured growth factor per step, predicted RK4(lambda dt))."""
    sizes = [1, 1, 1]
    sizes[axis] = n
    g = mod.make_grid(tuple(sizes), 2)
    h = float(g.axis("d" + "xyz"[axis])[0])
    xs = (np.arange(n) + 0.5) * h
    k = 2 * np.pi * mode
    phase = np.exp(1j * k * xs)
    q, field = init(g, phase)
    p = mod.Params(cfl_number=cfl, **av)
    mod.ghost_fill(g, p, q)
    dt = 0.8 * cfl * h  # fixed step
    inter = _interior(g)
    c0 = mode_coeff(q[field][inter], axis, mode, n)
    for _ in range(steps):
        mod.update(g, p, dt, q)
    c1 = mode_coeff(q[field][inter], axis, mode, n)
    lam = lam_fn(k_eff(k, h))
    return (c1 / c0), rk4_poly(lam * dt) ** steps


def _sound_init(axis, amp=1e-7, gm=5.0 / 3.0):
    ro0, p0 = 1.0, 1.0 / gm
    cs = math.sqrt(gm * p0 / ro0)

    def init(g, phase):
        q = uniform(g, ro=ro0, ei=p0 / (gm - 1) / ro0)
        shape = [1, 1, 1]
        shape[axis] = -1
        pert = amp * phase.real.reshape(shape)
        inter = _interior(g)
        q[0][inter] = ro0 + pert
        q[1 + axis][inter] = cs / ro0 * pert
        q[7][inter] = (p0 + cs * cs * pert) / (gm - 1) / (ro0 + pert)
        return q, 0
    return init, cs


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_linear_sound_wave(axis):
    init, cs = _sound_init(axis)
    got, want = _wave_run(axis, 32, 3, init, lambda ke: -1j * ke * cs)
    assert abs(got / want - 1) < 2e-6, (got, want)


def _alfven_init(axis, amp=1e-7):
    b0 = 0.8
    va = b0 / math.sqrt(4 * math.pi)
    t = (axis + 1) % 3  # transverse component

    def init(g, phase):
        b = [0.0, 0.0, 0.0]
        b[axis] = b0
        q = uniform(g, b=tuple(b), ei=1.5)
        shape = [1, 1, 1]
        shape[axis] = -1
        pert = amp * phase.real.reshape(shape)
        inter = _interior(g)
        q[4 + t][inter] = pert
        q[1 + t][inter] = -pert / math.sqrt(4 * math.pi)
        return q, 4 + t
    return init, va


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_linear_alfven_wave(axis):
    init, va = _alfven_init(axis)
    got, want = _wave_run(axis, 32, 2, init, lambda ke: -1j * ke * va)
    assert abs(got / want - 1) < 2e-6, (got, want)


def _fast_init(axis, amp=1e-7, gm=5.0 / 3.0):
    ro0, p0, b0 = 1.0, 0.6, 1.1
    t = (axis + 1) % 3
    cs2, va2 = gm * p0 / ro0, b0 * b0 / (4 * math.pi * ro0)
    vf = math.sqrt(cs2 + va2)

    def init(g, phase):
        b = [0.0, 0.0, 0.0]
        b[t] = b0
        q = uniform(g, ro=ro0, b=tuple(b), ei=p0 / (gm - 1) / ro0)
        shape = [1, 1, 1]
        shape[axis] = -1
        pert = amp * phase.real.reshape(shape)
        inter = _interior(g)
        q[0][inter] = ro0 + pert
        q[1 + axis][inter] = vf / ro0 * pert
        q[4 + t][inter] = b0 + b0 / ro0 * pert
        q[7][inter] = (p0 + cs2 * pert) / (gm - 1) / (ro0 + pert)
        return q, 0
    return init, vf


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_linear_fast_wave(axis):
    init, vf = _fast_init(axis)
    got, want = _wave_run(axis, 32, 2, init, lambda ke: -1j * ke * vf)
    assert abs(got / want - 1) < 2e-6, (got, want)


def _glm_run(axis, mod=O, n=32, mode=2, amp=1e-7, cfl=0.4, steps=4):
    """Divergence cleaning: a longitudinal field perturbation with v = 0
    evolves as dB/dt = -D psi, psi_s = damp_s (psi_0 - dt_s ch^2 D B_{s-1})
    stage by stage (mhd_integrator_host.hpp:238-272); the 2x2 stage
    recurrence in Fourier space is the prediction."""
    sizes = [1, 1, 1]
    sizes[axis] = n
    g = mod.make_grid(tuple(sizes), 2)
    h = float(g.axis("d" + "xyz"[axis])[0])
    xs = (np.arange(n) + 0.5) * h
    k = 2 * np.pi * mode
    q = uniform(g, ei=1.5)
    shape = [1, 1, 1]
    shape[axis] = -1
    inter = _interior(g)
    q[4 + axis][inter] = amp * np.cos(k * xs).reshape(shape)
    p = mod.Params(cfl_number=cfl, **NOAV)
    mod.ghost_fill(g, p, q)
    dt = 0.8 * cfl * h
    ch = 0.8 * cfl * g.min_dxyz / dt
    tau = 2.0 * dt
    D = 1j * k_eff(k, h)
    b = mode_coeff(q[4 + axis][inter], axis, mode, n)
    ps = 0.0 + 0j
    for _ in range(steps):
        mod.update(g, p, dt, q)
        b0, p0 = b, ps
        bs, pss = b0, p0
        for dts in (dt / 4.0, dt / 3.0, dt / 2.0, dt):
            bs, pss = b0 - dts * D * pss, (p0 - dts * ch * ch * D * bs) * math.exp(-dts / tau)
        b, ps = bs, pss
    got_b = mode_coeff(q[4 + axis][inter], axis, mode, n)
    got_p = mode_coeff(q[8][inter], axis, mode, n)
    return got_b, b, got_p, ps


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_divergence_cleaning_wave(axis):
    gb, wb, gp, wp = _glm_run(axis)
    assert abs(gb - wb) <= 1e-6 * abs(wb), (gb, wb)
    assert abs(gp - wp) <= 1e-6 * abs(wp), (gp, wp)


def _conservation(mod=O, av=None):
    from miso_amd import problems
    from miso_amd.mhd import Grid
    n = 12
    g = mod.make_grid((n, n, n), 2)
    pg = Grid(n, n, n, 2, device="cpu")
    q = problems.to_numpy(problems.turbulence(pg, v_rms=0.6, b_rms=0.5))
    p = mod.Params(cfl_number=0.4, **(av or {}))
    mod.ghost_fill(g, p, q)
    inter = (slice(None),) + _interior(g)

    def totals(q):
        ro, v, b, ei = q[0], q[1:4], q[4:7], q[7]
        et = ro * ei + 0.5 * ro * (v * v).sum(0) + (b * b).sum(0) / (8 * math.pi)
        return np.array([ro.sum(), *(ro * v).sum((1, 2, 3)), *b.sum((1, 2, 3)), et.sum()])

    qi = q[inter]
    scale = np.array([qi[0].sum()] + [np.abs(qi[0] * qi[1 + c]).sum() for c in range(3)] +
                     [np.abs(qi[4 + c]).sum() for c in range(3)] + [0.0])
    t0 = totals(qi)
    scale[7] = abs(t0[7])
    mod.run(g, p, q, 3)
    t1 = totals(q[inter])
    return np.abs(t1 - t0) / scale


def test_conservation():
    err = _conservation()
    assert err.max() < 1e-13, err


# ------------------------------------------------------ CFL brute force
def cfl_brute(g, q, gm, cfl):
    """min over interior cells of cfl*min(dx,dy,dz)/(cs+|v|+ca), every min
    taken on float-rounded values (utility.hpp:91-124)."""
    inter = _interior(g)
    ro, vx, vy, vz, bx, by, bz, ei = (q[m][inter] for m in range(8))
    cs = np.sqrt(gm * (gm - 1) * ei)
    vv = np.sqrt(vx * vx + vy * vy + vz * vz)
    ca = np.sqrt((bx * bx + by * by + bz * bz) / ro * (1.0 / (4 * math.pi)))
    hmin = np.minimum(np.minimum(g.axis("dx")[2:-2, None, None].astype(np.float32),
                                 g.axis("dy")[None, 2:-2, None].astype(np.float32)),
                      g.axis("dz")[None, None, 2:-2].astype(np.float32)).astype(np.float64)
    cand = cfl * hmin / (cs + vv + ca)
    return float(np.float32(min(cand.min(), 1e10)))


def test_cfl_brute_force():
    from tests.parity_util import random_state
    g = O.make_grid((9, 7, 6), 2, bounds=((0, 1), (0, 0.7), (0, 0.9)))
    q = random_state(g.shape, seed=5)
    assert O.cfl(g, O.Params(cfl_number=0.4), q) == cfl_brute(g, q, 5.0 / 3.0, 0.4)


def test_grid_matches_grid_hpp():
    g = O.make_grid((8, 1, 5), 2, bounds=((-0.5, 0.5), (0, 1), (0, 2)))
    assert g.shape == (12, 1, 9)
    x = g.axis("x")
    assert np.allclose(x, -0.5 + (np.arange(12) - 2 + 0.5) / 8, rtol=0, atol=1e-15)
    assert g.axis("dy")[0] == 1e30 and g.axis("dyi")[0] == 0.0
    assert g.min_dxyz == float(np.float32(min(1 / 8, 1e30, 2 / 5)))


# ------------------------------------- AV characteristic speed (A7)
CC_FACS = (1.3, 0.7, 2.1)  # cs_fac, vv_fac, ca_fac: distinct and non-zero


def char_vel_err(mod=O):
    """Largest relative miss of the oracle's frozen AV speed against the
    textbook speeds it combines (mhd_artificial_viscosity_host.hpp:57-64):
    the adiabatic sound speed sqrt(gamma p / rho) with p = (gamma-1) rho e,
    the flow speed |v| and the Gaussian-unit Alfven speed |B| / sqrt(4 pi rho).
    Each factor is switched on alone first, so a term paired with the wrong
    factor, a dropped term or a wrong constant all show."""
    from tests.parity_util import random_state
    g = mod.make_grid((6, 5, 4), 2)
    q = random_state(g.shape, seed=11, amp=0.8)
    gm = 5.0 / 3.0
    ro, v, b, e = q[0], q[1:4], q[4:7], q[7]
    pres = (gm - 1.0) * ro * e
    speeds = (np.sqrt(gm * pres / ro), np.sqrt((v * v).sum(0)),
              np.sqrt((b * b).sum(0)) / np.sqrt(4.0 * math.pi * ro))
    worst = 0.0
    cases = [tuple(f if n == m else 0.0 for n, f in enumerate(CC_FACS)) for m in range(3)]
    cases.append(CC_FACS)
    for cs_f, vv_f, ca_f in cases:
        p = mod.Params(gm=gm, cs_fac=cs_f, vv_fac=vv_f, ca_fac=ca_f)
        got = mod.char_vel(g, p, q)
        want = cs_f * speeds[0] + vv_f * speeds[1] + ca_f * speeds[2]
        worst = max(worst, float(np.max(np.abs(got - want) / want)))
    return worst


def test_char_vel_is_the_sum_of_textbook_speeds():
    assert char_vel_err() < 1e-14


# ------------------------------------- radiative transfer closed forms (A15)
def rt_linear_source_err(mod=O, kappa=3.0, a=0.8, b=1.7, i0=0.25, n=32):
    """One +z ray (w = 1) through a uniform absorber whose source function is
    linear in optical depth, S = a + b tau, entering with I0 at the lower
    face.  The transfer equation dI/dtau = S - I then has the exact solution
    I = I0 e^-tau + a (1 - e^-tau) + b (tau - 1 + e^-tau), and the linear-
    source short-characteristics integral (rt_host.hpp:667-678) is exact for
    S linear in tau, so the solver must reproduce it to rounding.  Q_rad is
    minus the divergence of the radiative flux F = 4 pi I z (the one ray
    stands for the full solid angle, 4 pi): Q = -4 pi kappa dI/dtau, with the
    derivative of the exact solution.  Returns (I error, Q error), relative
    to the largest value."""
    g = mod.make_grid((1, 1, n), 2, bounds=((-0.5, 0.5),) * 3)
    dz = g.axis("dz")[0]
    kap = np.full(g.shape, kappa)
    k = np.arange(g.shape[2]) - 1.0          # the lattice's lower face is cell 1
    tau_cell = kappa * dz * k
    S = np.broadcast_to(a + b * tau_cell, g.shape).copy()
    I = np.zeros((1,) + g.shape)
    mod.rt_solve(g, kap, S, I, periodic=(0, 0, 0), face_kind=(0, 0, 0, 0, 3, 0),
                 face_value=(0, 0, 0, 0, i0, 0), tolerance=1e-15, max_iters=3, num_rays=1)
    tau = tau_cell[1:n + 2]
    ex = np.exp(-tau)
    want_i = i0 * ex + a * (1.0 - ex) + b * (tau - 1.0 + ex)
    got_i = I[0, 0, 0, 1:n + 2]
    err_i = float(np.max(np.abs(got_i - want_i)) / np.max(np.abs(want_i)))
    q = mod.rt_qrad(g, kap, S, I, num_rays=1)[0, 0, 1:n + 2]
    dI = -i0 * ex + a * ex + b * (1.0 - ex)
    want_q = -4.0 * math.pi * kappa * dI
    err_q = float(np.max(np.abs(q - want_q)) / np.max(np.abs(want_q)))
    return err_i, err_q


def rt_linear_field_err(mod=O):
    """All 24 Carlson A4 rays through a uniform absorber many optical depths
    thick whose source function is linear in position, S = S0 + g.x, with
    thermal inflow (I = S) on every face.  The transfer equation
    mu.grad I = kappa (S - I) has the particular solution
    I_r = S - (mu_r . g) / kappa, and the inflow's departure from it decays as
    exp(-tau) along the ray; bilinear interpolation on the upwind face is
    exact for linear fields and the segment integral is exact for S linear in
    tau, so nodes more than ~40 optical depths from every face must hold the
    particular solution.  The solver takes the direction cosines through
    util::max2, which rounds them to float (rt_host.hpp:462-466), so the
    segment it integrates over points along mu rounded to float and the
    offset (mu_r . g) / kappa is reproduced to float precision: the miss is
    measured relative to that offset.  The cells are anisotropic
    (dx != dy != dz), so the two transverse interpolation weights of every
    ray differ and a transposed bilinear corner shows.  Returns the largest
    miss over the offset."""
    sizes = (48, 44, 40)
    g = mod.make_grid(sizes, 2, bounds=((0.0, 1.2), (0.0, 1.32), (0.0, 1.4)))
    kappa = 100.0                          # at least 2.5 optical depths per segment
    grad = np.array([0.3, -0.2, 0.45])
    x, y, z = g.axis("x"), g.axis("y"), g.axis("z")
    S = 5.0 + grad[0] * x[:, None, None] + grad[1] * y[None, :, None] + grad[2] * z[None, None, :]
    kap = np.full(g.shape, kappa)
    I = np.zeros((24,) + g.shape)
    mod.rt_solve(g, kap, S, I, periodic=(0, 0, 0), face_kind=(1,) * 6, tolerance=1e-15,
                 max_iters=3)
    w, mx, my, mz = mod.rt_quadrature(24)
    # lattice nodes 1 .. n+1 (cells); keep those >= 17 cells from every face
    sl = tuple(slice(1 + 17, 1 + s - 17) for s in sizes)
    worst = 0.0
    for r in range(24):
        off = (mx[r] * grad[0] + my[r] * grad[1] + mz[r] * grad[2]) / kappa
        worst = max(worst, float(np.max(np.abs(I[r][sl] - (S[sl] - off))) / abs(off)))
    return worst


def test_rt_linear_source_in_tau_is_exact():
    err_i, err_q = rt_linear_source_err()
    assert err_i < 1e-13, err_i
    assert err_q < 1e-11, err_q


def test_rt_linear_field_gives_the_particular_solution():
    assert rt_linear_field_err() < 1e-6


# ------------------------------------------------ source terms (A4, f2)
def source_stage_err(mod=O):
    """One RK stage from a uniform fluid at rest with the point-mass source
    (src_kind 2: src.v = -ro g x_a / r^3, src.ei = h ro; the functor of
    demo/mhd3d_magnetosphere/src/src.hpp:18-34 plus a heating).  Every flux
    divergence of a uniform state is exactly zero, so update_sc4
    (mhd_integrator_host.hpp:121-243) reduces to ro' = ro, v' = dt src.v / ro
    and ei' = (ro ei - ro v'^2 / 2 + src.ei) / ro: the momentum source inside
    dt, the heating outside it.  Returns the largest relative miss."""
    g = mod.make_grid((6, 5, 4), 2, bounds=((1.0, 2.0), (-0.5, 0.5), (0.5, 1.5)))
    gg, heat, dt, ro0, ei0 = 0.3, 0.05, 1e-3, 1.3, 1.5
    q = uniform(g, ro=ro0, ei=ei0)
    out = np.zeros_like(q)
    p = mod.Params(src_kind=2, grav=gg, src_heat=heat)
    mod.rk_stage(g, p, dt, 1.0, 2.0 * dt, q, q, out)
    inner = _interior(g)
    x, y, z = g.axis("x"), g.axis("y"), g.axis("z")
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    r3 = np.sqrt(X * X + Y * Y + Z * Z) ** 3
    v = [dt * (-ro0 * gg * c / r3) / ro0 for c in (X, Y, Z)]
    ke = 0.5 * ro0 * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
    ei = (ro0 * ei0 - ke + heat * ro0) / ro0
    worst = float(np.max(np.abs(out[0][inner] - ro0)) / ro0)
    for m, want in ((1, v[0]), (2, v[1]), (3, v[2]), (7, ei)):
        worst = max(worst, float(np.max(np.abs(out[m][inner] - want[inner])) /
                                 np.max(np.abs(want[inner]))))
    return worst


def test_source_terms_enter_the_stage_where_the_reference_puts_them():
    assert source_stage_err() < 1e-13


# --------------------------------------------------- planted mistakes
PLANTS = {
    1: "pressure-gradient sign in the x momentum equation",
    2: "magnetic pressure dropped from the x momentum equation",
    3: "RK stage 2 step dt/2 instead of dt/3",
    4: "stencil weight 7 instead of 8",
    5: "pressure dropped from the enthalpy of the energy flux",
    6: "limiter picks the steepest slope",
    7: "periodic halo shifted by one cell",
    8: "sign of the div-B term in the psi equation",
    9: "Alfven tension factor 1/8pi instead of 1/4pi",
    10: "magnetic energy dropped from the AV total energy",
    11: "Alfven speed dropped from the CFL speed",
    12: "min/max without float narrowing",
    13: "Alfven speed dropped from the frozen AV speed",
    14: "RT upwind and downwind source weights swapped",
    15: "RT bilinear corners u10 and u01 transposed",
    16: "Q_rad prefactor 2 pi instead of 4 pi",
    17: "source heating src.ei multiplied by dt",
}


def _pins_fail(mod):
    """Run the cheap pins against a variant; return the names that fail."""
    fails = []
    init, cs = _sound_init(0)
    got, want = _wave_run(0, 32, 3, init, lambda ke: -1j * ke * cs, mod=mod)
    if abs(got / want - 1) >= 2e-6:
        fails.append("sound")
    init, vf = _fast_init(0)
    got, want = _wave_run(0, 32, 2, init, lambda ke: -1j * ke * vf, mod=mod)
    if abs(got / want - 1) >= 2e-6:
        fails.append("fast")
    init, va = _alfven_init(0)
    got, want = _wave_run(0, 32, 2, init, lambda ke: -1j * ke * va, mod=mod)
    if abs(got / want - 1) >= 2e-6:
        fails.append("alfven")
    gb, wb, gp, wp = _glm_run(0, mod=mod)
    if abs(gb - wb) > 1e-6 * abs(wb) or abs(gp - wp) > 1e-6 * abs(wp):
        fails.append("glm")
    err = _conservation(mod=mod)
    if err.max() >= 1e-13:
        fails.append("conservation")
    # limiter against the generalised-minmod definition (and, when built,
    # the reference's own limiter.hpp)
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b, c = rng.standard_normal(3)
        if mod.dqq_eval(a, b, c, 1.5) != minmod_slope(a, b, c, 1.5):
            fails.append("limiter")
            break
    # CFL against the brute force
    from tests.parity_util import random_state
    g = mod.make_grid((9, 7, 6), 2, bounds=((0, 1), (0, 0.7), (0, 0.9)))
    q = random_state(g.shape, seed=5)
    if mod.cfl(g, mod.Params(cfl_number=0.4), q) != cfl_brute(g, q, 5.0 / 3.0, 0.4):
        fails.append("cfl")
    if char_vel_err(mod) >= 1e-14:
        fails.append("char_vel")
    err_i, err_q = rt_linear_source_err(mod)
    if err_i >= 1e-13:
        fails.append("rt_linear_source")
    if err_q >= 1e-11:
        fails.append("rt_qrad")
    if rt_linear_field_err(mod) >= 1e-6:
        fails.append("rt_linear_field")
    if source_stage_err(mod) >= 1e-13:
        fails.append("source_stage")
    return fails


@pytest.mark.parametrize("plant", list(PLANTS))
def test_planted_mistakes_are_caught(plant):
    ref_cfl_lib = O.lib()  # make sure the good build is loaded first
    assert ref_cfl_lib is not None
    with planted(plant) as mod:
        fails = _pins_fail(mod)
    assert fails, f"plant {plant} ({PLANTS[plant]}) passed every pin"
