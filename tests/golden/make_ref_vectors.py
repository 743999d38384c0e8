"""Writes tests/golden/ref_limiter_quadrature.npz from oracle/_ref (the
reference's own limiter.hpp, utility.hpp, constants.hpp and
rt_quadrature.hpp compiled from their sources by oracle/Makefile).

The file holds inputs and the reference's outputs only: dqq_eval and
flux_core on seeded arguments (FP64 and FP32), the float-narrowing min/max
on a few pairs, the constants, and the 24-ray Carlson A4 quadrature.
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
    np.savez_compressed(path, **out)
    print("wrote", path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
