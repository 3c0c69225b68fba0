"""ctypes binding of ``libswfd.so`` (the C ABI declared in ``include/swfd.h``).

The library is built in-tree by :mod:`soundwavesimulation_amd.build`. There is
no fallback: if the library is missing or no HIP device is visible, every
device entry point raises.
"""

from __future__ import annotations

import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
# SWFD_LIBRARY selects another build of the same library (e.g. a tuning variant)
LIBRARY = os.environ.get("SWFD_LIBRARY") or os.path.join(HERE, "libswfd.so")

SWFD_OK = 0
SWFD_EINVAL = 1
SWFD_EHIP = 2

_c_int = ctypes.c_int32
_c_float = ctypes.c_float
_fp = ctypes.c_void_p  # device pointers travel as plain addresses


class Grid(ctypes.Structure):
    _fields_ = [("ny", _c_int), ("nx", _c_int), ("inv_dx2", _c_float), ("inv_dy2", _c_float),
                ("center", _c_float), ("dt2", _c_float)]


class Medium(ctypes.Structure):
    _fields_ = [("c2", _fp), ("c2dt2", _fp), ("mur", _fp)]


class Receivers(ctypes.Structure):
    _fields_ = [("n", _c_int), ("row0", _fp), ("col0", _fp), ("weights", _fp),
                ("n_entries", _c_int), ("cell_map", _fp), ("entry_start", _fp),
                ("entry_cell", _fp), ("pair_rec", _fp), ("pair_w", _fp),
                ("footprint_rows", _fp)]


class StepArgs(ctypes.Structure):
    _fields_ = [("n_shots", _c_int), ("adjoint", _c_int), ("u_old", _fp), ("u", _fp),
                ("u_next", _fp), ("u_next_copy", _fp), ("src_rows", _fp), ("src_cols", _fp),
                ("src_values", _fp), ("src_index", _c_int), ("dense", _fp), ("residual", _fp),
                ("image_field", _fp), ("grad_in", _fp), ("grad_out", _fp),
                ("image_scale", _c_float), ("samples", _fp), ("src_weights", _fp)]


class ForwardArgs(ctypes.Structure):
    _fields_ = [("n_shots", _c_int), ("step_begin", _c_int), ("n_steps", _c_int),
                ("init_old", _fp), ("init_u", _fp), ("work0", _fp), ("work1", _fp),
                ("work2", _fp), ("work3", _fp), ("final_old", _fp), ("final_u", _fp), ("src_rows", _fp),
                ("src_cols", _fp), ("src_values", _fp), ("samples", _fp), ("hist", _fp),
                ("hist_first", _c_int), ("hist_every", _c_int), ("hist_slots", _c_int),
                ("ckpt", _fp), ("ckpt_every", _c_int), ("ckpt_slots", _c_int),
                ("hist_lap", _c_int), ("work4", _fp), ("work5", _fp), ("src_weights", _fp),
                ("bnd", _fp), ("bnd_first", _c_int), ("bnd_slots", _c_int)]


class AdjointArgs(ctypes.Structure):
    _fields_ = [("n_shots", _c_int), ("step_hi", _c_int), ("step_lo", _c_int),
                ("init_old", _fp), ("init_u", _fp), ("work0", _fp), ("work1", _fp),
                ("work2", _fp), ("work3", _fp), ("final_old", _fp), ("final_u", _fp), ("residual", _fp),
                ("hist", _fp), ("hist_first", _c_int), ("hist_every", _c_int),
                ("hist_slots", _c_int), ("image_scale", _c_float), ("grad", _fp),
                ("workspace", _fp), ("hist_lap", _c_int),
                ("recon_next", _fp), ("recon_cur", _fp), ("recon_work0", _fp),
                ("recon_work1", _fp), ("recon_work2", _fp), ("recon_work3", _fp),
                ("bnd", _fp), ("bnd_first", _c_int), ("bnd_slots", _c_int),
                ("src_rows", _fp), ("src_cols", _fp), ("src_values", _fp),
                ("src_weights", _fp)]


#: every symbol include/swfd.h declares
EXPORTS = (
    "swfd_last_error", "swfd_version", "swfd_medium_floats", "swfd_prepare_medium",
    "swfd_medium_view", "swfd_step", "swfd_gather", "swfd_scatter", "swfd_laplacian",
    "swfd_forward", "swfd_adjoint", "swfd_adjoint_workspace_floats",
    "swfd_residual_workspace_floats", "swfd_residual", "swfd_forward_batch_table_bytes",
    "swfd_forward_batch_prepare", "swfd_forward_batch", "swfd_adjoint_batch_table_bytes",
    "swfd_adjoint_batch_prepare", "swfd_adjoint_batch", "swfd_tune", "swfd_tune_clear",
    "swfd_lap_history", "swfd_steps_per_launch", "swfd_footprint_rows",
    "swfd_residual_noalloc", "swfd_history_convert", "swfd_adjoint_launch_shape",
    "swfd_boundary_floats", "swfd_adjoint_recon",
)

#: storage codes of swfd_history_convert (include/swfd.h)
HISTORY_F16 = 1
HISTORY_BF16 = 2

#: partial sums the built-in scratch of swfd_residual_noalloc holds (include/swfd.h)
RESIDUAL_SCRATCH_FLOATS = 1 << 20

_lib = None


def load_library(path=LIBRARY):
    """Load ``libswfd.so`` and declare argument types (no device is touched)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m soundwavesimulation_amd.build` "
            "(hipcc, gfx950). There is no CPU fallback."
        )
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.swfd_last_error.restype = ctypes.c_char_p
    lib.swfd_version.restype = ctypes.c_char_p
    lib.swfd_medium_floats.restype = ctypes.c_size_t
    lib.swfd_medium_floats.argtypes = [_c_int, _c_int]
    lib.swfd_prepare_medium.argtypes = [P(Grid), vp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, vp, vp]
    lib.swfd_medium_view.argtypes = [P(Grid), vp, P(Medium)]
    lib.swfd_footprint_rows.argtypes = [_c_int, _c_int, _c_int, vp, vp, vp]
    lib.swfd_medium_view.restype = None
    lib.swfd_step.argtypes = [P(Grid), P(Medium), P(Receivers), P(StepArgs), vp]
    lib.swfd_gather.argtypes = [P(Grid), P(Receivers), _c_int, vp, vp, vp]
    lib.swfd_scatter.argtypes = [P(Grid), P(Receivers), _c_int, vp, vp, vp]
    lib.swfd_laplacian.argtypes = [P(Grid), _c_int, vp, vp, vp]
    lib.swfd_forward.argtypes = [P(Grid), P(Medium), P(Receivers), P(ForwardArgs), vp]
    lib.swfd_adjoint.argtypes = [P(Grid), P(Medium), P(Receivers), P(AdjointArgs), vp]
    lib.swfd_adjoint_workspace_floats.restype = ctypes.c_size_t
    lib.swfd_adjoint_workspace_floats.argtypes = [_c_int, _c_int, _c_int]
    lib.swfd_residual_workspace_floats.restype = ctypes.c_size_t
    lib.swfd_residual_workspace_floats.argtypes = [_c_int, _c_int, _c_int]
    lib.swfd_residual.argtypes = [_c_int, _c_int, _c_int, vp, vp, ctypes.c_int64,
                                  ctypes.c_int64, vp, vp, vp, vp, vp, vp]
    lib.swfd_residual_noalloc.argtypes = [_c_int, _c_int, _c_int, vp, vp, ctypes.c_int64,
                                          ctypes.c_int64, vp, vp, vp, vp, vp]
    lib.swfd_history_convert.argtypes = [ctypes.c_size_t, _c_int, _c_int, vp, vp, vp]
    size_p = P(ctypes.c_size_t)
    for kind, args_t in (("forward", ForwardArgs), ("adjoint", AdjointArgs)):
        getattr(lib, f"swfd_{kind}_batch_table_bytes").argtypes = [
            _c_int, P(Grid), P(Medium), P(Receivers), P(args_t), size_p]
        getattr(lib, f"swfd_{kind}_batch_prepare").argtypes = [
            _c_int, P(Grid), P(Medium), P(Receivers), P(args_t), vp, ctypes.c_size_t]
        getattr(lib, f"swfd_{kind}_batch").argtypes = [
            _c_int, P(Grid), P(Medium), P(Receivers), P(args_t), vp, ctypes.c_size_t, vp]
    lib.swfd_tune.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int]
    lib.swfd_tune_clear.restype = None
    lib.swfd_tune_clear.argtypes = []
    lib.swfd_lap_history.restype = _c_int
    lib.swfd_lap_history.argtypes = [_c_int, _c_int]
    lib.swfd_steps_per_launch.restype = _c_int
    lib.swfd_steps_per_launch.argtypes = [_c_int, _c_int, _c_int, _c_int]
    lib.swfd_adjoint_launch_shape.argtypes = [_c_int, _c_int, _c_int, _c_int, vp]
    lib.swfd_boundary_floats.restype = ctypes.c_size_t
    lib.swfd_boundary_floats.argtypes = [_c_int, _c_int]
    lib.swfd_adjoint_recon.restype = _c_int
    lib.swfd_adjoint_recon.argtypes = [_c_int, _c_int]
    for name in ("swfd_prepare_medium", "swfd_step", "swfd_gather", "swfd_scatter",
                 "swfd_laplacian", "swfd_forward", "swfd_adjoint", "swfd_residual",
                 "swfd_residual_noalloc", "swfd_history_convert", "swfd_adjoint_launch_shape",
                 "swfd_forward_batch_table_bytes", "swfd_forward_batch_prepare",
                 "swfd_forward_batch", "swfd_adjoint_batch_table_bytes",
                 "swfd_adjoint_batch_prepare", "swfd_adjoint_batch", "swfd_tune",
                 "swfd_footprint_rows"):
        getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def lib():
    return load_library()


def check(rc):
    """Turn a swfd status into the exception the reference raises."""
    if rc == SWFD_OK:
        return
    message = lib().swfd_last_error().decode(errors="replace")
    if rc == SWFD_EINVAL:
        raise ValueError(message)
    raise RuntimeError(f"HIP error in libswfd: {message}")


def ptr(tensor):
    """Device address of a tensor (None -> NULL)."""
    if tensor is None:
        return None
    return ctypes.c_void_p(tensor.data_ptr())


def stream_of(device):
    """Handle of torch's current HIP stream on ``device``."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_device(device):
    """Validate that ``device`` is a HIP device this build can drive."""
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError(
            f"The MI355X backend runs on a HIP device (torch 'cuda'); got {device}. "
            "There is no CPU path."
        )
    if not torch.cuda.is_available():
        raise RuntimeError("No HIP device is visible to PyTorch.")
    load_library()
    return torch.device("cuda", torch.cuda.current_device() if device.index is None
                        else device.index)


def adjoint_launch_shape(ny, nx, n_shots, recon=True):
    """``{"steps_per_launch", "full_width", "segment_rows", "segments", "shot_groups",
    "shots_per_group"}`` of the imaging adjoint runs that rebuild the forward
    field (``recon``) or read a history (needs the device)."""
    out = (ctypes.c_int32 * 6)()
    check(lib().swfd_adjoint_launch_shape(ny, nx, n_shots, 1 if recon else 0, out))
    keys = ("steps_per_launch", "full_width", "segment_rows", "segments", "shot_groups",
            "shots_per_group")
    return dict(zip(keys, (int(v) for v in out)))
