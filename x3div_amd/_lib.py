"""ctypes binding of libx3dhip.so (the C ABI declared in include/x3dhip.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libx3dhip.so")

X3D_OK = 0
X3D_NPASS = 10
ERRORS = {
    -1: "X3D_ERR_ARG", -2: "X3D_ERR_SIZE", -3: "X3D_ERR_HIP", -4: "X3D_ERR_FFT",
    -5: "X3D_ERR_ALLOC", -6: "X3D_ERR_UNSUPPORTED",
}

# every symbol include/x3dhip.h declares, with its ctypes signature
_vp, _dp, _i, _d, _sz = C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_size_t
_ip = C.POINTER(C.c_int)
_hp = C.POINTER(C.c_double)
SIGNATURES = {
    "x3d_version": (C.c_int, []),
    "x3d_strerror": (C.c_char_p, [_i]),
    "x3d_create": (_i, [_i, _i, _i, _d, _d, _d, C.POINTER(_vp)]),
    "x3d_destroy": (_i, [_vp]),
    "x3d_set_stream": (_i, [_vp, _vp]),
    "x3d_sync": (_i, [_vp]),
    "x3d_wtime": (_d, []),
    "x3d_get_sizes": (_i, [_vp, _ip, _ip, _ip]),
    "x3d_field_alloc": (_i, [_vp, _sz, C.POINTER(_vp)]),
    "x3d_field_free": (_i, [_vp, _vp]),
    "x3d_memcpy_h2d": (_i, [_vp, _vp, _vp, _sz]),
    "x3d_memcpy_d2h": (_i, [_vp, _vp, _vp, _sz]),
    "x3d_operator_get": (_i, [_vp, _i, _hp, _hp, _hp, _hp, _hp, _ip, _ip]),
    "x3d_operator_set": (_i, [_vp, _i, _hp, _hp, _hp, _hp, _d]),
    "x3d_scheme_consts": (_i, [_vp, _hp]),
    "x3d_xthomas": (_i, [_vp, _dp, _i, _i, _i, _i]),
    "x3d_ythomas": (_i, [_vp, _dp, _i, _i, _i, _i]),
    "x3d_zthomas": (_i, [_vp, _dp, _i, _i, _i, _i]),
    "x3d_derx": (_i, [_vp, _dp, _dp, _i, _i, _i, _i]),
    "x3d_dery": (_i, [_vp, _dp, _dp, _i, _dp, _i, _i, _i]),
    "x3d_derz": (_i, [_vp, _dp, _dp, _i, _i, _i, _i]),
    "x3d_derxx": (_i, [_vp, _dp, _dp, _i, _i, _i, _i]),
    "x3d_deryy": (_i, [_vp, _dp, _dp, _i, _i, _i, _i]),
    "x3d_derzz": (_i, [_vp, _dp, _dp, _i, _i, _i, _i]),
    "x3d_derxvp": (_i, [_vp, _dp, _dp, _i, _i, _i, _i, _i]),
    "x3d_interxvp": (_i, [_vp, _dp, _dp, _i, _i, _i, _i, _i]),
    "x3d_derxpv": (_i, [_vp, _dp, _dp, _i, _i, _i, _i, _i]),
    "x3d_interxpv": (_i, [_vp, _dp, _dp, _i, _i, _i, _i, _i]),
    "x3d_interyvp": (_i, [_vp, _dp, _dp, _i, _i, _i, _i, _i]),
    "x3d_deryvp": (_i, [_vp, _dp, _dp, _i, _dp, _i, _i, _i, _i]),
    "x3d_interypv": (_i, [_vp, _dp, _dp, _i, _i, _i, _i, _i]),
    "x3d_derypv": (_i, [_vp, _dp, _dp, _i, _dp, _i, _i, _i, _i]),
    "x3d_derzvp": (_i, [_vp, _dp, _dp, _i, _i, _i, _i, _i]),
    "x3d_interzvp": (_i, [_vp, _dp, _dp, _i, _i, _i, _i, _i]),
    "x3d_derzpv": (_i, [_vp, _dp, _dp, _i, _i, _i, _i, _i]),
    "x3d_interzpv": (_i, [_vp, _dp, _dp, _i, _i, _i, _i, _i]),
    "x3d_transpose_x_to_y": (_i, [_vp, _vp, _vp, _sz, _i, _i, _i]),
    "x3d_transpose_y_to_z": (_i, [_vp, _vp, _vp, _sz, _i, _i, _i]),
    "x3d_transpose_z_to_y": (_i, [_vp, _vp, _vp, _sz, _i, _i, _i]),
    "x3d_transpose_y_to_x": (_i, [_vp, _vp, _vp, _sz, _i, _i, _i]),
    "x3d_poisson": (_i, [_vp, _dp]),
    "x3d_poisson_tables": (_i, [_vp, _hp, _hp, _hp, _hp, _hp, _hp, _hp]),
    "x3d_momentum_rhs": (_i, [_vp] + [_dp] * 6),
    "x3d_int_time": (_i, [_vp, _d] + [_dp] * 6),
    "x3d_divergence": (_i, [_vp] + [_dp] * 4),
    "x3d_gradp": (_i, [_vp] + [_dp] * 4),
    "x3d_cor_vel": (_i, [_vp] + [_dp] * 6),
    "x3d_solve_poisson": (_i, [_vp] + [_dp] * 7),
    "x3d_step": (_i, [_vp, _d] + [_dp] * 10),
    "x3d_step_noalloc": (_i, [_vp, _d] + [_dp] * 10),
    "x3d_graph_begin": (_i, [_vp]),
    "x3d_graph_end": (_i, [_vp]),
    "x3d_fill": (_i, [_vp, _dp, C.c_size_t, _d]),
    "x3d_xthomas_tab": (_i, [_vp, _dp, _hp, _hp, _hp, _hp, _d, _i, _i, _i]),
    "x3d_ythomas_tab": (_i, [_vp, _dp, _hp, _hp, _hp, _hp, _d, _i, _i, _i]),
    "x3d_zthomas_tab": (_i, [_vp, _dp, _hp, _hp, _hp, _hp, _d, _i, _i, _i]),
    "x3d_thomas1d": (_i, [_hp, _hp, _hp, _hp, _i]),
    "x3d_step_timed": (_i, [_vp, _d] + [_dp] * 10 + [C.POINTER(C.c_float)]),
    "x3d_derx_f32": (_i, [_vp, _dp, _dp, _i, _i, _i, _i]),
    "x3d_dery_f32": (_i, [_vp, _dp, _dp, _i, _dp, _i, _i, _i]),
    "x3d_derz_f32": (_i, [_vp, _dp, _dp, _i, _i, _i, _i]),
    "x3d_apply_f32": (_i, [_vp, _i, _dp, _dp]),
    "x3d_poisson_f32": (_i, [_vp, _dp]),
    "x3d_apply_circ_mfma": (_i, [_vp, _i, _dp, _dp]),
    "x3d_step_f32": (_i, [_vp, _d] + [_dp] * 10),
    "x3d_step_timed_f32": (_i, [_vp, _d] + [_dp] * 10 + [C.POINTER(C.c_float)]),
    "x3d_create_bc": (_i, [_i, _i, _i, _d, _d, _d, _ip, _ip, C.POINTER(_vp)]),
    "x3d_set_stretching": (_i, [_vp, _i, _vp, _vp]),
    "x3d_bfp_encode": (_i, [_vp, _dp, _dp]),
    "x3d_bfp_decode": (_i, [_vp, _dp, _dp]),
    "x3d_apply_bfp": (_i, [_vp, _i, _dp, _dp]),
    "x3d_step_bfp": (_i, [_vp, _d] + [_dp] * 10),
    "x3d_step_timed_bfp": (_i, [_vp, _d] + [_dp] * 10 + [C.POINTER(C.c_float)]),
    "x3d_graph_capture_step": (_i, [_vp, _d] + [_dp] * 10),
    "x3d_graph_capture_step_f32": (_i, [_vp, _d] + [_dp] * 10),
    "x3d_graph_launch": (_i, [_vp]),
    "x3d_batch_create": (_i, [_i, C.POINTER(C.c_int), C.POINTER(C.c_double), _i, C.POINTER(_vp)]),
    "x3d_batch_destroy": (_i, [_vp]),
    "x3d_batch_set_stream": (_i, [_vp, _vp]),
    "x3d_batch_sizes": (_i, [_vp, _ip, C.POINTER(C.c_longlong)]),
    "x3d_batch_field": (_i, [_vp, _i, _i, C.POINTER(_vp)]),
    "x3d_batch_context": (_i, [_vp, _i, C.POINTER(_vp)]),
    "x3d_batch_init_fields": (_i, [_vp, _i, C.c_uint64, _d]),
    "x3d_batch_step": (_i, [_vp, _d]),
    "x3d_batch_graph_capture": (_i, [_vp, _d]),
    "x3d_batch_graph_launch": (_i, [_vp]),
    "x3d_batch_autotune": (_i, [_vp, _i, C.c_char_p, _sz]),
    "x3d_pass_name": (C.c_char_p, [_i]),
    "x3d_pass_bytes_per_point": (_d, [_i]),
    "x3d_test_err": (_i, [_vp, _dp, _i, _hp]),
    "x3d_div_stats": (_i, [_vp, _dp, _hp, _hp]),
    "x3d_init_fields": (_i, [_vp, _i, C.c_uint64, _d, _dp, _dp, _dp]),
}

_lib = None


def load(path: str | None = None):
    """Load libx3dhip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python -m x3div_amd.build). "
            "x3div_amd has no CPU fallback.")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class X3DError(RuntimeError):
    def __init__(self, fn: str, code: int):
        msg = load().x3d_strerror(code).decode()
        super().__init__(f"{fn} failed: {ERRORS.get(code, code)} ({msg})")
        self.code = code


def check(fn: str, code: int) -> None:
    if code != X3D_OK:
        raise X3DError(fn, code)
