"""ctypes binding of ``libswno.so`` (declarations in ``include/swno.h``).

The library must be built (``python -m soundwavesimulation_amd.build``);
there is no fallback: loading raises if it is missing, and every entry point
raises ``ValueError`` (bad argument) or ``RuntimeError`` (HIP error) with the
library's message when it fails.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_lib = None

_i32, _i64, _f32, _vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_void_p


class GemmArgs(ctypes.Structure):
    _fields_ = [("M", _i32), ("N", _i32), ("K", _i32), ("batch", _i32), ("k_split", _i32),
                ("A", _vp), ("a_rs", _i64), ("a_cs", _i64), ("a_bs", _i64),
                ("B", _vp), ("b_rs", _i64), ("b_cs", _i64), ("b_bs", _i64),
                ("C", _vp), ("c_rs", _i64), ("c_cs", _i64), ("c_bs", _i64),
                ("alpha", _f32), ("beta", _f32)]


#: every symbol include/swno.h declares
EXPORTS = ("swno_last_error", "swno_gemm_workspace_floats", "swno_gemm", "swno_bias_act",
           "swno_gelu_backward", "swno_col_sum_workspace_floats", "swno_col_sum",
           "swno_weight_grad_workspace_floats", "swno_weight_grad", "swno_head_forward",
           "swno_head_backward_workspace_floats", "swno_head_backward", "swno_layer_forward",
           "swno_layer_backward", "swno_x_transform",
           "swno_spectral_pack", "swno_mix_assemble", "swno_mix_assemble_backward",
           "swno_mse_workspace_floats", "swno_mse", "swno_adamw")


def library_path():
    return os.environ.get("SWNO_LIBRARY", os.path.join(_HERE, "libswno.so"))


def load_library():
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with "
                           "`python -m soundwavesimulation_amd.build` (no CPU fallback).")
    lib = ctypes.CDLL(path)
    lib.swno_last_error.restype = ctypes.c_char_p
    lib.swno_gemm_workspace_floats.restype = ctypes.c_size_t
    lib.swno_gemm_workspace_floats.argtypes = [ctypes.POINTER(GemmArgs)]
    lib.swno_gemm.argtypes = [ctypes.POINTER(GemmArgs), _vp, _vp]
    lib.swno_bias_act.argtypes = [_i64, _i32, _vp, _vp, _vp, _vp, _i32, _i32, _vp]
    lib.swno_gelu_backward.argtypes = [_i64, _vp, _vp, _vp, _vp]
    lib.swno_col_sum_workspace_floats.restype = ctypes.c_size_t
    lib.swno_col_sum_workspace_floats.argtypes = [_i64, _i32]
    lib.swno_col_sum.argtypes = [_i64, _i32, _vp, _f32, _vp, _vp, _vp]
    lib.swno_weight_grad_workspace_floats.restype = ctypes.c_size_t
    lib.swno_weight_grad_workspace_floats.argtypes = [_i64, _i32, _i32]
    lib.swno_weight_grad.argtypes = [_i64, _i32, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _vp,
                                     _i32, _vp]
    lib.swno_head_forward.argtypes = [_i64, _i32, _i32] + [_vp] * 7 + [_i32, _vp]
    lib.swno_head_backward_workspace_floats.restype = ctypes.c_size_t
    lib.swno_head_backward_workspace_floats.argtypes = [_i64]
    lib.swno_head_backward.argtypes = [_i64, _i32, _i32] + [_vp] * 11 + [_i32, _vp]
    lib.swno_layer_forward.argtypes = [_i32, _i32, _i32, _i32] + [_vp] * 7 + [_i32, _i32, _vp]
    lib.swno_layer_backward.argtypes = [_i32, _i32, _i32, _i32] + [_vp] * 7
    lib.swno_x_transform.argtypes = [_i32, _i32, _i32, _i32, _vp, _i64, _i64, _vp, _vp, _i32,
                                     _vp]
    lib.swno_spectral_pack.argtypes = [_i32, _i32, _i32, _i32, _vp, _vp, _vp]
    lib.swno_mix_assemble.argtypes = [_i32, _i32, _vp, _vp, _vp, _vp]
    lib.swno_mix_assemble_backward.argtypes = [_i32, _i32, _vp, _vp, _vp, _i32, _vp]
    lib.swno_mse_workspace_floats.restype = ctypes.c_size_t
    lib.swno_mse_workspace_floats.argtypes = [_i64]
    lib.swno_mse.argtypes = [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    lib.swno_adamw.argtypes = [_i64, _vp, _vp, _vp, _vp, _f32, _f32, _f32, _f32, _f32, _i32,
                               _vp]
    for name in EXPORTS:
        if not hasattr(lib, name):
            raise RuntimeError(f"{path} does not export {name}")
    _lib = lib
    return lib


def check(rc):
    if rc == 0:
        return
    msg = load_library().swno_last_error().decode()
    if rc == 1:
        raise ValueError(msg)
    raise RuntimeError(msg)
