"""Writes tests/golden/ref_limiter_quadrature.npz and
tests/golden/ref_table_interp.npz from oracle/_ref (the reference's own
limiter.hpp, utility.hpp, constants.hpp, rt_quadrature.hpp and
table_interpolator.hpp compiled from their sources by oracle/Makefile).

The files hold inputs and the reference's outputs only: dqq_eval and
flux_core on seeded arguments (FP64 and FP32), the float-narrowing min/max
on a few pairs, the constants, the 24-ray Carlson A4 quadrature, and 1D/2D
uniform-table interpolation of seeded tables at seeded points (inside, on
the nodes and outside the table range).
Run:  python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def main():
    from tests.test_oracle_pins import lib_ref
    L = lib_ref()
    rng = np.random.default_rng(20261202)
    vals = np.concatenate([rng.standard_normal(400), rng.standard_normal(200) * 1e-6,
                           [0.0, -0.0, 1e-30, -1e-30, 1.0, 1.0 + 1e-12]])
    out = {}
    for S, dt in (("f64", np.float64), ("f32", np.float32)):
        n = 4000
        dq_in = np.stack([rng.choice(vals, n), rng.choice(vals, n), rng.choice(vals, n),
                          rng.choice([1.0, 1.5, 2.0], n)], 1).astype(dt)
        fl_in = np.stack([rng.choice(vals, n), rng.choice(vals, n), rng.choice(vals, n),
                          rng.choice(vals, n), np.abs(rng.standard_normal(n)) + 0.1,
                          rng.choice([0.0, 0.5, 1.0], n)], 1).astype(dt)
        f_dq = getattr(L, f"ref_dqq_eval_{S}")
        f_fl = getattr(L, f"ref_flux_core_{S}")
        out[f"dqq_in_{S}"] = dq_in
        out[f"dqq_out_{S}"] = np.array([f_dq(*map(dt, r)) for r in dq_in], dtype=dt)
        out[f"flux_in_{S}"] = fl_in
        out[f"flux_out_{S}"] = np.array([f_fl(*map(dt, r)) for r in fl_in], dtype=dt)
        out[f"const_{S}"] = np.array([getattr(L, f"ref_const_{S}")(w) for w in range(4)], dtype=dt)
    mm = np.stack([rng.standard_normal(200), rng.standard_normal(200), rng.standard_normal(200)], 1)
    mm[0] = (1.0 + 1e-12, 2.0, 3.0)
    out["minmax_in"] = mm
    out["min2_out"] = np.array([L.ref_min2_f64(a, b) for a, b, _ in mm])
    out["min3_out"] = np.array([L.ref_min3_f64(a, b, c) for a, b, c in mm])
    out["max3_out"] = np.array([L.ref_max3_f64(a, b, c) for a, b, c in mm])
    L.ref_quadrature_f64.argtypes = [C.c_int, C.POINTER(C.c_double)]
    L.ref_quadrature_f64.restype = C.c_int
    q = np.zeros(96)
    nr = L.ref_quadrature_f64(24, q.ctypes.data_as(C.POINTER(C.c_double)))
    out["quadrature_24"] = q[:4 * nr].reshape(nr, 4)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_limiter_quadrature.npz")
    if not os.path.exists(path):  # already committed; its vectors do not change
        np.savez_compressed(path, **out)
        print("wrote", path, {k: v.shape for k, v in out.items()})
    tables(L)


def tables(L):
    """Uniform-table interpolation vectors (interpolate_uniform_table1d/2d)."""
    rng = np.random.default_rng(7070)
    out = {}
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    for S, dt, ct in (("f64", np.float64, C.c_double), ("f32", np.float32, C.c_float)):
        n = 37
        t1 = rng.standard_normal(n).astype(dt)
        lo, hi = dt(-1.25), dt(3.5)
        x = np.concatenate([rng.uniform(-2.0, 4.5, 3000), np.linspace(-1.25, 3.5, n),
                            [-1.25, 3.5, -100.0, 100.0]]).astype(dt)
        y = np.zeros_like(x)
        f = getattr(L, f"ref_table1d_{S}")
        f.argtypes = [C.POINTER(ct), C.c_int, ct, ct, C.POINTER(ct), C.POINTER(ct), C.c_int]
        f(P(t1, ct), n, lo, hi, P(x, ct), P(y, ct), len(x))
        out.update({f"t1_{S}": t1, "n1": np.array(n), f"lohi1_{S}": np.array([lo, hi], dt),
                    f"x1_{S}": x, f"y1_{S}": y})
        n0, n1 = 13, 21
        t2 = rng.standard_normal((n0, n1)).astype(dt)
        a0, b0, a1, b1 = dt(0.5), dt(2.0), dt(-3.0), dt(1.0)
        m = 4000
        x0 = np.concatenate([rng.uniform(0.0, 2.5, m), [0.5, 2.0, -9.0, 9.0]]).astype(dt)
        x1 = np.concatenate([rng.uniform(-3.5, 1.5, m), [-3.0, 1.0, 9.0, -9.0]]).astype(dt)
        y2 = np.zeros_like(x0)
        f = getattr(L, f"ref_table2d_{S}")
        f.argtypes = [C.POINTER(ct), C.c_int, C.c_int, ct, ct, ct, ct, C.POINTER(ct),
                      C.POINTER(ct), C.POINTER(ct), C.c_int]
        f(P(np.ascontiguousarray(t2), ct), n0, n1, a0, b0, a1, b1, P(x0, ct), P(x1, ct),
          P(y2, ct), len(x0))
        out.update({f"t2_{S}": t2, f"ranges2_{S}": np.array([a0, b0, a1, b1], dt),
                    f"x20_{S}": x0, f"x21_{S}": x1, f"y2_{S}": y2})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_table_interp.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
