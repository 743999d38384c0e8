"""The C++ host layer (include/miso_hip.hpp): MISO-shaped classes over the C
ABI.  Its tests are compiled into tests/cpp/_build/libmiso_cpp_tests.so and
called in this process through ctypes."""
import ctypes as C
import os

import numpy as np
import pytest

from tests.parity_util import assert_bitwise

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build",
                   "libmiso_cpp_tests.so")


def _lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB)
    L.cpp_test_hd1d_shock_tube.argtypes = [C.c_char_p, C.c_int]
    L.cpp_test_hd1d_shock_tube.restype = C.c_int
    L.cpp_run_steps.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_double, C.c_double,
                                C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                C.POINTER(C.c_double)]
    L.cpp_run_steps.restype = C.c_int
    return L


def test_cpp_test_library_exports():
    L = _lib()
    assert L.cpp_test_hd1d_shock_tube and L.cpp_run_steps


@pytest.mark.gpu
def test_cpp_hd1d_shock_tube():
    """miso/tests/serial/test_hd1d_shock_tube.cpp through the C++ layer:
    L1 < 5e-2 against the exact Riemann solution and x/y/z agreement."""
    buf = C.create_string_buffer(256)
    rc = _lib().cpp_test_hd1d_shock_tube(buf, 256)
    print(buf.value.decode(), flush=True)
    assert rc == 0, buf.value.decode()


CASES = {
    "orszag_tang_32_f64": dict(sizes=(32, 32, 1), steps=10, dtype=np.float64),
    "turbulence_12_f32": dict(sizes=(12, 12, 12), steps=4, dtype=np.float32),
    "stratified_fixed_z_8x8x64_f64": dict(sizes=(8, 8, 64), steps=3, dtype=np.float64,
                                          strat=True),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_cpp_layer_matches_oracle(case):
    from miso_amd import problems
    from miso_amd.mhd import Grid
    from miso_amd.rmhd import Stratification, stratified_box
    from oracle import oracle as O
    c = CASES[case]
    sizes, n, dt_ = c["sizes"], c["steps"], c["dtype"]
    if c.get("strat"):
        bounds = ((0.0, 0.125), (0.0, 0.125), (0.0, 1.0))
        kinds, src = (1, 1, 2), dict(src_kind=1, grav_axis=2, grav=Stratification().grav)
        g = Grid(*sizes, 2, bounds=bounds, device="cpu")
        q0 = problems.to_numpy(stratified_box(g, Stratification(), seed=2))
    else:
        bounds = ((0.0, 1.0),) * 3
        kinds, src = (1, 1, 1), dict(src_kind=0, grav_axis=2, grav=0.0)
        g = Grid(*sizes, 2, bounds=bounds, device="cpu")
        gen = problems.orszag_tang if sizes[2] == 1 else problems.turbulence
        q0 = problems.to_numpy(gen(g))
    q0 = np.ascontiguousarray(q0.astype(dt_))
    og = O.make_grid(sizes, 2, bounds=bounds, dtype=dt_)
    ref = q0.copy()
    p = O.Params(cfl_number=0.4, periodic=kinds, **src,
                 **({"fixed": q0.copy()} if 2 in kinds else {}))
    O.ghost_fill(og, p, ref)
    ref_dts = O.run(og, p, ref, n)
    out = np.empty_like(q0)
    dts = np.zeros(n)
    b = (C.c_double * 6)(*[v for lohi in bounds for v in lohi])
    k = (C.c_int * 3)(*kinds)
    s = (C.c_int * 27)(*([1] * 27))
    rc = _lib().cpp_run_steps(1 if dt_ == np.float64 else 0, *sizes, b, k, s, 5.0 / 3.0, 0.4,
                              src["src_kind"], src["grav_axis"], src["grav"], n,
                              q0.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                              dts.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    if dt_ == np.float64:
        assert np.array_equal(dts, np.asarray(ref_dts)), (dts, ref_dts)
        sl = (slice(None),) + g.interior()
        assert_bitwise(out[sl], ref[sl], case)
    else:
        # the FP32 oracle build and the GPU round differently (see the FP32
        # step test); the layer is held to the same 2e-4 bar there
        sl = (slice(None),) + g.interior()
        err = np.abs(out[sl].astype(np.float64) - ref[sl]).max(axis=(1, 2, 3)) / \
            np.maximum(np.abs(ref[sl]).max(axis=(1, 2, 3)), 1e-30)
        assert err.max() < 2e-4, err
