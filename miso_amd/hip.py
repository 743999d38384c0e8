"""ctypes binding of the C ABI in include/misohip.h (libmisohip.so).

This is the Python side of the drop-in boundary: plain pointers, sizes and
POD descriptors go across, exactly what a C/C++ caller would pass.  Device
memory is plumbing (torch tensors); the product is the HIP library.  A missing
library is an error -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmisohip.so")

F64, F32 = 0, 1

FIELD_NAMES = ("ro", "vx", "vy", "vz", "bx", "by", "bz", "ei", "ph")


class MisohipGrid(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "i_size", "j_size", "k_size", "i_total", "j_total", "k_total",
        "i_margin", "j_margin", "k_margin", "is_", "js", "ks")] + [
        ("min_dxyz", C.c_double),
        ("dx", C.c_void_p), ("dy", C.c_void_p), ("dz", C.c_void_p),
        ("dxi", C.c_void_p), ("dyi", C.c_void_p), ("dzi", C.c_void_p)]


class MisohipFields(C.Structure):
    _fields_ = [("q", C.c_void_p * 9)]


class MisohipBc(C.Structure):
    _fields_ = [("periodic", C.c_int * 3), ("sym_sign", (C.c_int * 9) * 3),
                ("fixed", C.c_void_p * 9)]


class MisohipParams(C.Structure):
    _fields_ = [("gm", C.c_double), ("cfl_number", C.c_double), ("ep", C.c_double),
                ("fh", C.c_double), ("cs_fac", C.c_double), ("ca_fac", C.c_double),
                ("vv_fac", C.c_double), ("src_kind", C.c_int), ("grav_axis", C.c_int),
                ("grav", C.c_double)]


class MisohipTiles(C.Structure):
    _fields_ = [("rk_pencil_bx", C.c_int), ("rk_pencil_by", C.c_int), ("sweep_chunk", C.c_int),
                ("rk_variant", C.c_int), ("rk_zchunk", C.c_int), ("av_fuse_xy", C.c_int),
                ("av_xy_chunk", C.c_int)]


class MisohipRtConfig(C.Structure):
    _fields_ = [("num_rays", C.c_int), ("periodic", C.c_int * 3), ("face_kind", C.c_int * 6),
                ("face_value", C.c_double * 6), ("tolerance", C.c_double), ("max_iters", C.c_int),
                ("x_rays_yfast", C.c_int), ("sweep_planes", C.c_int)]


class MisohipWork(C.Structure):
    _fields_ = [("argm", MisohipFields), ("rslt", MisohipFields), ("cc", C.c_void_p)]


# int (*)(void *user, const misohip_fields *q, void *stream)
BC_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(MisohipFields), C.c_void_p)
SRC_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(MisohipFields), C.POINTER(C.c_void_p),
                     C.c_void_p)

# misohip_memcpy directions
MEMCPY_H2D, MEMCPY_D2H, MEMCPY_D2D = 1, 2, 3
E_ARG = -2


# (name, argtypes) of every exported entry point of include/misohip.h
_P = C.POINTER
SIGNATURES = {
    "misohip_create": [C.c_int, C.c_void_p, _P(C.c_void_p)],
    "misohip_destroy": [C.c_void_p],
    "misohip_set_stream": [C.c_void_p, C.c_void_p],
    "misohip_synchronize": [C.c_void_p],
    "misohip_alloc": [C.c_void_p, C.c_size_t, _P(C.c_void_p)],
    "misohip_free": [C.c_void_p, C.c_void_p],
    "misohip_memcpy": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int],
    "misohip_strerror": [C.c_int],
    "misohip_copy_to_device": [C.c_void_p, _P(MisohipGrid), C.c_int, C.c_void_p, C.c_void_p],
    "misohip_copy_to_host": [C.c_void_p, _P(MisohipGrid), C.c_int, C.c_void_p, C.c_void_p],
    "misohip_ghost_fill": [C.c_void_p, _P(MisohipGrid), C.c_int, _P(MisohipFields),
                           _P(MisohipBc)],
    "misohip_rk_stage": [C.c_void_p, _P(MisohipGrid), C.c_int, _P(MisohipParams),
                         _P(MisohipFields), _P(MisohipFields), _P(MisohipFields), C.c_double,
                         C.c_double, C.c_double],
    "misohip_char_vel": [C.c_void_p, _P(MisohipGrid), C.c_int, _P(MisohipParams),
                         _P(MisohipFields), C.c_void_p],
    "misohip_av_sweep": [C.c_void_p, _P(MisohipGrid), C.c_int, _P(MisohipParams), C.c_int,
                         C.c_double, _P(MisohipFields), _P(MisohipFields), C.c_void_p],
    "misohip_cfl": [C.c_void_p, _P(MisohipGrid), C.c_int, _P(MisohipParams), _P(MisohipFields),
                    _P(C.c_double)],
    "misohip_update": [C.c_void_p, _P(MisohipGrid), C.c_int, _P(MisohipParams), _P(MisohipBc),
                       _P(MisohipFields), _P(MisohipWork), C.c_double],
    "misohip_step": [C.c_void_p, _P(MisohipGrid), C.c_int, _P(MisohipParams), _P(MisohipBc),
                     _P(MisohipFields), _P(MisohipWork), _P(C.c_double)],
    "misohip_stream_bench": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_size_t],
    "misohip_update_noalloc": [C.c_void_p, _P(MisohipGrid), C.c_int, _P(MisohipParams),
                               _P(MisohipBc), _P(MisohipFields), C.c_double],
    "misohip_table1d": [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_double,
                        C.c_void_p, C.c_void_p, C.c_long],
    "misohip_table2d": [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                        C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long],
    "misohip_update_heated": [C.c_void_p, _P(MisohipGrid), C.c_int, _P(MisohipParams), _P(MisohipBc),
                              _P(MisohipFields), _P(MisohipWork), C.c_double, C.c_void_p,
                              _P(C.c_double)],
    "misohip_set_bc_callback": [C.c_void_p, BC_FN, C.c_void_p],
    "misohip_get_bc_callback": [C.c_void_p, _P(C.c_void_p), _P(C.c_void_p)],
    "misohip_set_source_callback": [C.c_void_p, SRC_FN, C.c_void_p, _P(C.c_void_p)],
    "misohip_get_source_callback": [C.c_void_p, _P(C.c_void_p), _P(C.c_void_p),
                                    _P(C.c_void_p)],
    "misohip_rt_quadrature": [C.c_int, _P(C.c_double), _P(C.c_double), _P(C.c_double),
                              _P(C.c_double)],
    "misohip_rt_solve": [C.c_void_p, _P(MisohipGrid), _P(MisohipRtConfig), C.c_void_p, C.c_void_p,
                         C.c_void_p, C.c_void_p, _P(C.c_int), _P(C.c_double)],
    "misohip_rt_qrad": [C.c_void_p, _P(MisohipGrid), _P(MisohipRtConfig), C.c_void_p, C.c_void_p,
                        C.c_void_p, C.c_void_p],
    "misohip_rt_ray_layout": [C.c_void_p, _P(MisohipGrid), _P(MisohipRtConfig), _P(C.c_uint)],
    "misohip_rt_source": [C.c_void_p, _P(MisohipGrid), _P(MisohipFields), C.c_double, C.c_double,
                          C.c_double,
                          C.c_void_p, C.c_void_p],
    "misohip_rt_heat": [C.c_void_p, _P(MisohipGrid), _P(MisohipFields), C.c_void_p, C.c_double],
    "misohip_mlp_create": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                           _P(C.c_void_p)],
    "misohip_mlp_destroy": [C.c_void_p],
    "misohip_mlp_set_layer": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
    "misohip_mlp_get_layer": [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p],
    "misohip_mlp_set_norm": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p],
    "misohip_mlp_forward": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p],
    "misohip_mlp_grad": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                         _P(C.c_double)],
    "misohip_mlp_adam": [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float,
                         C.c_float],
    "misohip_mlp_features": [C.c_void_p, C.c_void_p, _P(MisohipGrid), _P(MisohipFields),
                             C.c_double, C.c_void_p, C.c_void_p],
    "misohip_mlp_qrad": [C.c_void_p, C.c_void_p, _P(MisohipGrid), C.c_void_p, C.c_void_p,
                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "misohip_mlp_residual": [C.c_void_p, C.c_void_p, _P(MisohipGrid), C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_double, C.c_void_p],
    "misohip_run": [C.c_void_p, _P(MisohipGrid), C.c_int, _P(MisohipParams), _P(MisohipBc),
                    _P(MisohipFields), _P(MisohipWork), C.c_int, _P(C.c_double)],
    "misohip_batch_create": [C.c_void_p, C.c_int, C.c_int, _P(MisohipGrid), _P(MisohipFields),
                             _P(MisohipWork), _P(MisohipParams), _P(MisohipBc), _P(C.c_void_p)],
    "misohip_batch_destroy": [C.c_void_p],
    "misohip_batch_capture": [C.c_void_p, C.c_void_p, C.c_int],
    "misohip_batch_step": [C.c_void_p, C.c_void_p, C.c_int],
    "misohip_batch_set_chunks": [C.c_void_p, C.c_void_p, _P(C.c_int)],
    "misohip_batch_state": [C.c_void_p, C.c_void_p, _P(C.c_double), _P(C.c_double),
                            _P(C.c_long), _P(C.c_int)],
    "misohip_set_tiles": [C.c_void_p, _P(MisohipTiles)],
    "misohip_get_tiles": [C.c_void_p, _P(MisohipTiles)],
    "misohip_profile_enable": [C.c_void_p, C.c_int],
    "misohip_profile_read": [C.c_void_p, _P(C.c_double), _P(C.c_int), C.c_int],
}

KIND_NAMES = ("rk_stage", "av_x", "av_y", "av_z", "ghost_fill", "cfl", "rt", "mlp")

_lib = None


class MisohipError(RuntimeError):
    pass


def lib():
    """Load libmisohip.so; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MisohipError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (make -C miso_amd/csrc); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, argt in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = C.c_int
        L.misohip_strerror.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().misohip_strerror(rc).decode()
        raise MisohipError(f"{what}: {msg} (code {rc})")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
