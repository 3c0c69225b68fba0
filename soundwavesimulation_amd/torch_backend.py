"""Device backend of the 2D ring-array FDTD with the reference's call surface.

Drop-in counterpart of the reference module ``fdtd2d/torch_backend.py``: the
same classes, functions, argument names and error messages, with every
field update, gather and scatter running in the hand-written gfx950 kernels
of ``libswfd.so`` (C ABI in ``include/swfd.h``). PyTorch provides device
memory, streams and a few validation reductions; it runs none of the
stencil, boundary, source, gather or scatter arithmetic. There is no CPU
path: a non-HIP device or a missing library raises.

Precision: fields are float32 (the reference's GPU dtype); float64 is not
used on the device. ``dtype`` arguments other than ``torch.float32`` raise.

Buffer semantics: every single-step call returns a new field tensor, as the
reference's eager ops do, so fields a caller keeps are never overwritten.
Multi-step runs (``simulate_shot`` without snapshots, ``simulate_shots_batch``)
rotate three internal buffers inside one native call.
"""

from __future__ import annotations

import math
import os

import numpy as np
import torch

from . import _native
from .sources import linear_chirp as _linear_chirp_np

_F32 = torch.float32


# ---------------------------------------------------------------------------
# runtime helpers (torch_backend.py:17-77)
# ---------------------------------------------------------------------------

def resolve_device(requested: str = "auto") -> torch.device:
    """Resolve ``auto``/``cuda[:N]`` to a HIP device; other devices are refused."""
    requested = str(requested).lower()
    if requested == "auto":
        if torch.cuda.is_available():
            return torch.device("cuda", torch.cuda.current_device())
        raise ValueError("No HIP device is available; this backend has no CPU path.")
    try:
        device = torch.device(requested)
    except (RuntimeError, ValueError) as exc:
        raise ValueError(f"Unknown device {requested!r}; use auto, cuda, or cuda:N.") from exc
    if device.type != "cuda":
        raise ValueError("The MI355X FDTD backend supports only HIP (cuda) devices.")
    if not torch.cuda.is_available():
        raise ValueError("CUDA was requested, but PyTorch cannot access a CUDA GPU.")
    index = torch.cuda.current_device() if device.index is None else device.index
    if index < 0 or index >= torch.cuda.device_count():
        raise ValueError(f"CUDA device {index} is unavailable; found {torch.cuda.device_count()}.")
    return torch.device("cuda", index)


def configure_runtime(device: torch.device, cpu_threads: int = 0, allow_tf32: bool = False) -> None:
    """Host threads and current device (TF32 has no effect: no matmul on the path)."""
    if cpu_threads < 0:
        raise ValueError("cpu_threads must be zero (automatic) or positive.")
    if cpu_threads:
        torch.set_num_threads(cpu_threads)
    device = torch.device(device)
    if device.type == "cuda":
        torch.cuda.set_device(device)
        torch.backends.cuda.matmul.allow_tf32 = bool(allow_tf32)


def device_summary(device: torch.device) -> str:
    device = torch.device(device)
    if device.type == "cuda":
        props = torch.cuda.get_device_properties(device)
        return (f"HIP {device.index or 0}: {props.name}; {props.multi_processor_count} CUs; "
                f"{props.total_memory / 1024**3:.1f} GiB")
    return f"CPU; {torch.get_num_threads()} PyTorch threads"


def linear_chirp(dt, duration_s, f_start_hz, f_end_hz, *, device, dtype=_F32):
    """Hann-windowed linear chirp (float64 on the host, cast to ``dtype``)."""
    pulse = _linear_chirp_np(dt, duration_s, f_start_hz, f_end_hz)
    return torch.as_tensor(pulse, dtype=dtype, device=device)


def _check_dtype(dtype):
    if dtype not in (None, _F32):
        raise ValueError(f"The HIP kernels propagate float32 fields; dtype {dtype} is not supported.")


# ---------------------------------------------------------------------------
# model
# ---------------------------------------------------------------------------

class TorchScalarWave2D:
    """Speed model ``c`` (ny, nx) on one HIP device (torch_backend.py:105-174)."""

    def __init__(self, c, dx: float, dy: float | None = None, *, device, dtype=_F32):
        _check_dtype(dtype)
        self.device = _native.require_device(device)
        self.dtype = _F32
        self._c = torch.as_tensor(np.asarray(c) if not isinstance(c, torch.Tensor) else c,
                                  dtype=_F32).to(self.device).contiguous().clone()
        if self._c.ndim != 2:
            raise ValueError("Sound speed c must be a 2D array (ny, nx).")
        if not bool(torch.isfinite(self._c).all()) or not bool((self._c > 0).all()):
            raise ValueError("Sound speed must contain finite positive values.")
        self._dx = float(dx)
        self._dy = float(dx if dy is None else dy)
        if self._dx <= 0.0 or self._dy <= 0.0:
            raise ValueError("Grid spacing must be positive.")
        if min(self._c.shape) < 3:
            raise ValueError("The grid must be at least 3 x 3 cells.")
        self._c2 = self._c.square()
        self.version = 0
        inv_dx2 = 1.0 / self._dx**2
        inv_dy2 = 1.0 / self._dy**2
        self.laplacian_kernel = torch.tensor(
            [[0.0, inv_dy2, 0.0], [inv_dx2, -2.0 * (inv_dx2 + inv_dy2), inv_dx2],
             [0.0, inv_dy2, 0.0]], device=self.device, dtype=_F32).reshape(1, 1, 3, 3)
        self._grid = _native.Grid(ny=self.shape[0], nx=self.shape[1], inv_dx2=inv_dx2,
                                  inv_dy2=inv_dy2, center=-2.0 * (inv_dx2 + inv_dy2), dt2=0.0)

    @property
    def shape(self):
        return tuple(self._c.shape)

    @property
    def dx(self):
        return self._dx

    @property
    def dy(self):
        return self._dy

    @property
    def c_max(self):
        return float(self._c.max().item())

    @property
    def c(self):
        return self._c

    @property
    def c2(self):
        return self._c2

    def set_c(self, c):
        """Update the speed in place (same tensors), validating like the reference."""
        updated = torch.as_tensor(np.asarray(c) if not isinstance(c, torch.Tensor) else c,
                                  dtype=_F32).to(self.device)
        if tuple(updated.shape) != self.shape:
            raise ValueError(f"Sound speed shape {tuple(updated.shape)} does not match {self.shape}.")
        if not bool(torch.isfinite(updated).all()) or not bool((updated > 0).all()):
            raise ValueError("Sound speed must contain finite positive values.")
        self._c.copy_(updated)
        self._c2.copy_(updated.square())
        self.version += 1

    def laplacian(self, field):
        """Zero-padded 5-point Laplacian of a (ny, nx) or (S, ny, nx) field."""
        field = torch.as_tensor(field, device=self.device, dtype=_F32).contiguous()
        if tuple(field.shape[-2:]) != self.shape or field.ndim not in (2, 3):
            raise ValueError("Field must have shape (ny, nx) or (shots, ny, nx).")
        out = torch.empty_like(field)
        n = 1 if field.ndim == 2 else field.shape[0]
        _native.check(_native.lib().swfd_laplacian(
            self._grid, n, _native.ptr(field), _native.ptr(out), _native.stream_of(self.device)))
        return out


# ---------------------------------------------------------------------------
# single-shot solver
# ---------------------------------------------------------------------------

class TorchFDTD2D:
    """Leapfrog solver for one shot (torch_backend.py:236-344).

    ``compile_step`` is accepted for signature compatibility; the step is
    already one fused native kernel.
    """

    def __init__(self, model: TorchScalarWave2D, dt: float, compile_step=False):
        self.model = model
        self.dt = float(dt)
        self.dt2 = self.dt * self.dt
        dev = model.device
        self.u_old = torch.zeros(model.shape, device=dev, dtype=_F32)
        self.u = torch.zeros_like(self.u_old)
        self.source_row = 0
        self.source_col = 0
        self.compiled = bool(compile_step)
        self._grid = _native.Grid(ny=model._grid.ny, nx=model._grid.nx,
                                  inv_dx2=model._grid.inv_dx2, inv_dy2=model._grid.inv_dy2,
                                  center=model._grid.center, dt2=self.dt2)
        self._coef = torch.empty(int(_native.lib().swfd_medium_floats(*model.shape)),
                                 device=dev, dtype=_F32)
        self._medium = _native.Medium()
        _native.lib().swfd_medium_view(self._grid, _native.ptr(self._coef), self._medium)
        self._medium_version = None
        self._src_rc = torch.zeros(2, device=dev, dtype=torch.int32)
        self._src_w = None
        self._refresh_boundaries()
        if self.cfl >= 1.0:
            raise ValueError(f"Unstable time step: CFL = {self.cfl:.3f} >= 1.")

    # -- medium coefficients ------------------------------------------------
    def _refresh_boundaries(self):
        m = self.model
        _native.check(_native.lib().swfd_prepare_medium(
            self._grid, _native.ptr(m.c), self.dt, m.dx, m.dy, _native.ptr(self._coef),
            _native.stream_of(m.device)))
        self._medium_version = m.version

    def _sync_medium(self):
        if self._medium_version != self.model.version:
            self._refresh_boundaries()

    def _mur(self):
        ny, nx = self.model.shape
        base = 2 * ny * nx
        return self._coef[base:base + 2 * ny + 2 * nx + 8]

    @property
    def k_left(self):
        ny, _ = self.model.shape
        return self._mur()[1:ny - 1]

    @property
    def k_right(self):
        ny, _ = self.model.shape
        return self._mur()[ny + 1:2 * ny - 1]

    @property
    def k_bottom(self):
        ny, nx = self.model.shape
        return self._mur()[2 * ny + 1:2 * ny + nx - 1]

    @property
    def k_top(self):
        ny, nx = self.model.shape
        return self._mur()[2 * ny + nx + 1:2 * ny + 2 * nx - 1]

    @property
    def corner_kx(self):
        ny, nx = self.model.shape
        return self._mur()[2 * ny + 2 * nx:2 * ny + 2 * nx + 4]

    @property
    def corner_ky(self):
        ny, nx = self.model.shape
        return self._mur()[2 * ny + 2 * nx + 4:2 * ny + 2 * nx + 8]

    @property
    def cfl(self):
        return float(self.model.c_max * self.dt
                      * math.sqrt(1.0 / self.model.dx**2 + 1.0 / self.model.dy**2))

    def set_source(self, row: int, col: int):
        self.source_row = int(row)
        self.source_col = int(col)
        self._src_rc = torch.tensor([self.source_row, self.source_col],
                                    dtype=torch.int32).to(self.model.device)
        self._src_w = None

    def set_bilinear_source(self, row0: int, col0: int, weights):
        """Inject ``w * value`` on the 2 x 2 nodes from (row0, col0) with the
        weights w00, w10, w01, w11 (``RingArray.bilinear_sources``) instead of
        ``value`` at one node: source and receivers are then adjoint, which
        makes the simulation reciprocal (no counterpart in the reference's
        solver; its suggested fix, desmond/docs/upstream_suggestions.md:65-71)."""
        ny, nx = self.model.shape
        if not (0 <= int(row0) < ny - 1 and 0 <= int(col0) < nx - 1):
            raise ValueError("Bilinear source footprint must lie inside the grid.")
        self.set_source(row0, col0)
        self._src_w = _weights_tensor(weights, 1, self.model.device)

    def set_c(self, c):
        self.model.set_c(c)
        self._refresh_boundaries()
        if self.cfl >= 1.0:
            raise ValueError(f"Updated medium is unstable: CFL = {self.cfl:.3f} >= 1.")

    def reset(self):
        self.u_old = torch.zeros_like(self.u_old)
        self.u = torch.zeros_like(self.u)

    # -- stepping -------------------------------------------------------------
    def _value_tensor(self, value):
        if isinstance(value, torch.Tensor):
            t = value.reshape(-1)[:1]
            if t.device != self.model.device or t.dtype != _F32:
                t = t.to(device=self.model.device, dtype=_F32)
            return t.contiguous()
        return torch.tensor([float(value)], device=self.model.device, dtype=_F32)

    def _step(self, value, source_field, source_mode, adjoint=False):
        self._sync_medium()
        out = torch.empty_like(self.u)
        args = _native.StepArgs(n_shots=1, adjoint=int(bool(adjoint)),
                                u_old=_native.ptr(self.u_old), u=_native.ptr(self.u),
                                u_next=_native.ptr(out))
        keep = []
        if source_mode == 1:
            vt = self._value_tensor(value)
            keep.append(vt)
            args.src_rows = _native.ptr(self._src_rc[0:1])
            args.src_cols = _native.ptr(self._src_rc[1:2])
            args.src_values = _native.ptr(vt)
            args.src_index = 0
            args.src_weights = _native.ptr(self._src_w)
        elif source_mode == 2:
            args.dense = _native.ptr(source_field)
        _native.check(_native.lib().swfd_step(self._grid, self._medium, None, args,
                                              _native.stream_of(self.model.device)))
        self.u_old, self.u = self.u, out
        return self.u

    def inject_and_step(self, value):
        return self._step(value, None, 1, adjoint=False)

    def step(self, zero=None):
        return self._step(zero, None, 0, adjoint=False)

    def inject_field_and_step(self, source_field, adjoint=True):
        source_field = torch.as_tensor(source_field, device=self.model.device, dtype=_F32)
        if tuple(source_field.shape) != self.model.shape:
            raise ValueError("Dense source field must match the FDTD grid shape.")
        return self._step(None, source_field.contiguous(), 2, adjoint=adjoint)

    def _adjoint_image_step(self, source_field, forward_field, gradient, scale):
        """Adjoint step with a dense source, then ``gradient += scale * u' * L[forward_field]``
        in place in one native call (the reference's ``inject_field_and_step`` +
        ``laplacian`` + ``addcmul_``, invert_ring.py:668-681); returns the new field."""
        self._sync_medium()
        if tuple(source_field.shape) != self.model.shape or \
                tuple(forward_field.shape) != self.model.shape:
            raise ValueError("Dense source field must match the FDTD grid shape.")
        if gradient.dtype != _F32 or not gradient.is_contiguous() or \
                tuple(gradient.shape) != self.model.shape:
            raise ValueError("gradient must be a contiguous float32 (ny, nx) tensor")
        fwd = forward_field
        if fwd.dtype != _F32:
            fwd = convert_history(fwd.contiguous(), torch.empty(fwd.shape, device=fwd.device,
                                                                dtype=_F32))
        out = torch.empty_like(self.u)
        args = _native.StepArgs(n_shots=1, adjoint=1, u_old=_native.ptr(self.u_old),
                                u=_native.ptr(self.u), u_next=_native.ptr(out),
                                dense=_native.ptr(source_field.contiguous()),
                                image_field=_native.ptr(fwd.contiguous()),
                                grad_in=_native.ptr(gradient), grad_out=_native.ptr(gradient),
                                image_scale=float(scale))
        _native.check(_native.lib().swfd_step(self._grid, self._medium, None, args,
                                              _native.stream_of(self.model.device)))
        self.u_old, self.u = self.u, out
        return self.u


# ---------------------------------------------------------------------------
# receivers
# ---------------------------------------------------------------------------

def scatter_table(row0, col0, weights, shape):
    """Deterministic bilinear scatter table for the device kernels.

    Pairs (receiver, weight) are listed corner-major (w00, w10, w01, w11) and
    by receiver within a corner; each distinct footprint cell sums its pairs
    in that order. Returns (cell_map (ny*nx,), entry_start (E+1,),
    entry_cell (E,), pair_rec (P,), pair_w (P,)) as NumPy arrays.
    """
    ny, nx = shape
    row0 = np.asarray(row0, dtype=np.int64)
    col0 = np.asarray(col0, dtype=np.int64)
    w = np.asarray(weights, dtype=np.float64).reshape(4, -1)
    R = row0.size
    cells, recs, ws = [], [], []
    for k, (dr, dc) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
        cells.append((row0 + dr) * nx + (col0 + dc))
        recs.append(np.arange(R))
        ws.append(w[k])
    cells = np.concatenate(cells)
    recs = np.concatenate(recs)
    ws = np.concatenate(ws)
    order = np.argsort(cells, kind="stable")
    cells, recs, ws = cells[order], recs[order], ws[order]
    uniq, start = np.unique(cells, return_index=True)
    entry_start = np.append(start, cells.size).astype(np.int32)
    cell_map = np.full(ny * nx, -1, dtype=np.int32)
    cell_map[uniq] = np.arange(uniq.size, dtype=np.int32)
    return (cell_map, entry_start, uniq.astype(np.int32), recs.astype(np.int32),
            ws.astype(np.float32))


class TorchRingSampler:
    """Bilinear receiver gather / scatter on the device (torch_backend.py:347-440)."""

    def __init__(self, ring_array, *, device, dtype=_F32):
        _check_dtype(dtype)
        device = _native.require_device(device)
        self.device = device
        self.n_elements = ring_array.n_elements
        self.grid_shape = (ring_array._y_axis.size, ring_array._x_axis.size)
        for name in ("_row0", "_row1", "_col0", "_col1"):
            setattr(self, name, torch.as_tensor(getattr(ring_array, name), device=device,
                                                dtype=torch.long))
        for name in ("_w00", "_w10", "_w01", "_w11"):
            setattr(self, name, torch.as_tensor(getattr(ring_array, name), device=device,
                                                dtype=_F32))
        self.receivers = (self._row0, self._row1, self._col0, self._col1,
                          self._w00, self._w10, self._w01, self._w11)
        weights = np.stack([ring_array._w00, ring_array._w10, ring_array._w01, ring_array._w11])
        ny, nx = self.grid_shape
        r0 = np.asarray(ring_array._row0)
        c0 = np.asarray(ring_array._col0)
        if r0.min() < 0 or c0.min() < 0 or r0.max() + 1 >= ny or c0.max() + 1 >= nx:
            raise ValueError("Receiver footprints must lie inside the grid.")
        self._r0 = torch.as_tensor(ring_array._row0, dtype=torch.int32).to(device)
        self._c0 = torch.as_tensor(ring_array._col0, dtype=torch.int32).to(device)
        self._w = torch.as_tensor(weights, dtype=_F32).to(device).contiguous()
        table = scatter_table(ring_array._row0, ring_array._col0, weights, self.grid_shape)
        (self._cell_map, self._entry_start, self._entry_cell, self._pair_rec,
         self._pair_w) = (torch.as_tensor(a).to(device) for a in table)
        # rows and full-width wavefronts holding footprint cells: sampled forward
        # runs keep intermediate steps' fields only there (same samples)
        self._footprint_rows = None
        if os.environ.get("SWFD_FOOTPRINT_ROWS", "1") != "0":
            rows = np.zeros(ny, dtype=np.uint32)
            r0_h = np.ascontiguousarray(r0, dtype=np.int32)
            c0_h = np.ascontiguousarray(c0, dtype=np.int32)
            _native.check(_native.lib().swfd_footprint_rows(
                ny, nx, int(r0_h.size), r0_h.ctypes.data, c0_h.ctypes.data, rows.ctypes.data))
            self._footprint_rows = torch.as_tensor(rows.view(np.int32)).to(device)
        self._rec = _native.Receivers(
            n=self.n_elements, row0=_native.ptr(self._r0), col0=_native.ptr(self._c0),
            weights=_native.ptr(self._w), n_entries=int(self._entry_cell.numel()),
            cell_map=_native.ptr(self._cell_map), entry_start=_native.ptr(self._entry_start),
            entry_cell=_native.ptr(self._entry_cell), pair_rec=_native.ptr(self._pair_rec),
            pair_w=_native.ptr(self._pair_w),
            footprint_rows=_native.ptr(self._footprint_rows))
        ny, nx = self.grid_shape
        self._grid = _native.Grid(ny=ny, nx=nx, inv_dx2=0.0, inv_dy2=0.0, center=0.0, dt2=0.0)

    def _gather(self, fields, n_shots):
        out = torch.empty((n_shots, self.n_elements), device=self.device, dtype=_F32)
        _native.check(_native.lib().swfd_gather(self._grid, self._rec, n_shots,
                                                _native.ptr(fields), _native.ptr(out),
                                                _native.stream_of(self.device)))
        return out

    def record(self, field):
        field = torch.as_tensor(field, device=self.device, dtype=_F32)
        if tuple(field.shape) != self.grid_shape:
            raise ValueError("Field must have shape (ny, nx) matching the grid.")
        return self._gather(field.contiguous(), 1)[0]

    def record_batch(self, fields):
        """Record every receiver for each field in a ``(shots, ny, nx)`` batch."""
        fields = torch.as_tensor(fields, device=self.device, dtype=_F32)
        if fields.ndim != 3 or tuple(fields.shape[1:]) != self.grid_shape:
            raise ValueError("Batched fields must have shape (shots, ny, nx) matching the grid.")
        return self._gather(fields.contiguous(), fields.shape[0])

    def _scatter(self, residual, n_shots, out):
        shape = (n_shots, *self.grid_shape)
        if out is None or tuple(out.shape) != shape or not out.is_contiguous() \
                or out.dtype != _F32 or out.device != self.device:
            target = torch.empty(shape, device=self.device, dtype=_F32)
        else:
            target = out
        _native.check(_native.lib().swfd_scatter(self._grid, self._rec, n_shots,
                                                 _native.ptr(residual), _native.ptr(target),
                                                 _native.stream_of(self.device)))
        if out is not None and target is not out:
            out.copy_(target.reshape(out.shape))
            return out
        return target

    def record_adjoint_batch(self, residual, out=None):
        """Scatter batched residuals ``(shots, receivers)`` into ``(shots, ny, nx)``."""
        residual = torch.as_tensor(residual, device=self.device, dtype=_F32)
        if residual.ndim != 2 or residual.shape[1] != self.n_elements:
            raise ValueError("Batched residuals must have shape (shots, receivers).")
        return self._scatter(residual.contiguous(), residual.shape[0], out)

    def record_adjoint(self, residual, out=None):
        """Scatter receiver residuals; exact transpose of :meth:`record`."""
        residual = torch.as_tensor(residual, device=self.device, dtype=_F32)
        if tuple(residual.shape) != (self.n_elements,):
            raise ValueError(f"Residual shape {tuple(residual.shape)} does not match "
                             f"({self.n_elements},).")
        result = self._scatter(residual.contiguous(), 1,
                               None if out is None else out.reshape(1, *self.grid_shape))
        return result.reshape(self.grid_shape) if out is None else out


# ---------------------------------------------------------------------------
# multi-step runs
# ---------------------------------------------------------------------------

def _history_code(wave_dtype):
    """Storage code of a reduced-precision history dtype, None for float32."""
    if wave_dtype in (None, _F32):
        return None
    if wave_dtype == torch.float16:
        return _native.HISTORY_F16
    if wave_dtype == torch.bfloat16:
        return _native.HISTORY_BF16
    raise ValueError("wave_dtype must be float32, float16 or bfloat16 on the device.")


def convert_history(src, dst):
    """Pack float32 fields into a float16 / bfloat16 history or unpack them back
    (the reference's ``wave[k].copy_(fields)`` and ``u_fwd.to(dtype)``), rounding
    to nearest even, on the device."""
    if src.numel() != dst.numel() or not src.is_contiguous() or not dst.is_contiguous():
        raise ValueError("history conversion needs contiguous tensors of one size")
    if src.dtype == _F32:
        direction, code = 0, _history_code(dst.dtype)
    else:
        direction, code = 1, _history_code(src.dtype)
        if dst.dtype != _F32:
            raise ValueError("history conversion unpacks into float32")
    if code is None:
        raise ValueError("history conversion needs a float16 or bfloat16 side")
    _native.check(_native.lib().swfd_history_convert(
        src.numel(), direction, code, _native.ptr(src), _native.ptr(dst),
        _native.stream_of(src.device)))
    return dst


def _source_values_tensor(pulse, n_steps, device, source_values=None):
    if source_values is None:
        values = torch.zeros(n_steps, device=device, dtype=_F32)
        p = torch.as_tensor(pulse).reshape(-1)
        k = min(p.numel(), n_steps)
        values[:k].copy_(p[:k].to(device=device, dtype=_F32))
        return values
    source_values = torch.as_tensor(source_values)
    if tuple(source_values.shape) != (n_steps,):
        raise ValueError("source_values must contain one value per time step.")
    return source_values.to(device=device, dtype=_F32).contiguous()


def _check_src_weights(src_weights, n_shots):
    if src_weights is None:
        return None
    if src_weights.dtype != _F32 or not src_weights.is_contiguous() or \
            tuple(src_weights.shape) != (n_shots, 4):
        raise ValueError("Bilinear source weights must be a contiguous float32 (shots, 4) tensor.")
    return src_weights


def _weights_tensor(weights, n_shots, device):
    if weights is None:
        return None
    t = torch.as_tensor(np.asarray(weights, dtype=np.float32).reshape(n_shots, 4)).to(device)
    return t.contiguous()


def forward_run(solver, n_shots, rows_i32, cols_i32, values, n_steps, *, init=None,
                samples=None, sampler=None, hist=None, hist_first=0, hist_every=1, ckpt=None,
                ckpt_every=0, step_begin=0, work=None, final=True, hist_lap=False,
                src_weights=None, bnd=None, bnd_first=0):
    """Run forward steps in one native call; returns the final ``(u_old, u)``.

    ``init`` is the state before ``step_begin`` (zeros if None); ``samples``
    (N, S, R), ``hist`` (slots, S, ny, nx) and ``ckpt`` (slots, S, ny, nx) are
    optional outputs (see ``swfd_forward_args``). ``work`` holds three, four or
    six scratch fields; with four the run advances several steps per launch
    (with samples: two, four with six fields).
    ``hist_lap``: the history receives the Laplacians of the fields (of
    ``init`` and every output but the last) for ``adjoint_run(hist_lap=True)``.
    ``src_weights`` (S, 4) float32 device tensor: bilinear sources, the rows
    and columns then being the footprint corners (``RingArray.bilinear_sources``).
    ``bnd`` (slots, S, swfd_boundary_floats) float32: boundary ring, the
    boundary cells of the field after step j going to slot ``j - bnd_first``
    (full-width runs with more than two steps per launch need six work fields).
    """
    model = solver.model
    dev = model.device
    shape = (n_shots, *model.shape)
    if work is None:  # six fields with samples: four steps per launch (see swfd_forward_args)
        n_work = 6 if samples is not None else 4
        work = [torch.empty(shape, device=dev, dtype=_F32) for _ in range(n_work)]
    if init is None:
        work[0].zero_()
        work[1].zero_()
        init = (work[0], work[1])
    solver._sync_medium()
    args = _native.ForwardArgs(
        n_shots=n_shots, step_begin=step_begin, n_steps=n_steps,
        init_old=_native.ptr(init[0]), init_u=_native.ptr(init[1]),
        work0=_native.ptr(work[0]), work1=_native.ptr(work[1]), work2=_native.ptr(work[2]),
        work3=_native.ptr(work[3]) if len(work) > 3 else None,
        final_old=_native.ptr(work[0]) if final else None,
        final_u=_native.ptr(work[1]) if final else None,
        src_rows=_native.ptr(rows_i32), src_cols=_native.ptr(cols_i32),
        src_values=_native.ptr(values), samples=_native.ptr(samples),
        hist=_native.ptr(hist), hist_first=hist_first, hist_every=hist_every,
        hist_slots=0 if hist is None else hist.shape[0],
        ckpt=_native.ptr(ckpt), ckpt_every=ckpt_every,
        ckpt_slots=0 if ckpt is None else ckpt.shape[0], hist_lap=1 if hist_lap else 0,
        work4=_native.ptr(work[4]) if len(work) > 5 else None,
        work5=_native.ptr(work[5]) if len(work) > 5 else None,
        src_weights=_native.ptr(_check_src_weights(src_weights, n_shots)),
        bnd=_native.ptr(_check_ring(bnd, n_shots, model.shape)), bnd_first=bnd_first,
        bnd_slots=0 if bnd is None else bnd.shape[0])
    rec = None
    if samples is not None:
        if sampler is None:
            raise ValueError("samples need a sampler")
        if tuple(sampler.grid_shape) != tuple(model.shape):
            raise ValueError("Sampler grid does not match the solver grid.")
        rec = sampler._rec
    _native.check(_native.lib().swfd_forward(solver._grid, solver._medium, rec, args,
                                             _native.stream_of(dev)))
    return (work[0], work[1]) if final else None


def boundary_floats(shape):
    """Floats per shot in a slot of the boundary ring of a (ny, nx) grid."""
    return int(_native.lib().swfd_boundary_floats(int(shape[0]), int(shape[1])))


def _check_ring(bnd, n_shots, shape):
    if bnd is None:
        return None
    if (bnd.dtype != _F32 or bnd.dim() != 3 or bnd.shape[1] != n_shots
            or bnd.shape[2] != boundary_floats(shape) or not bnd.is_contiguous()):
        raise ValueError("boundary ring must be a contiguous float32 (slots, shots, "
                         "boundary_floats) tensor.")
    return bnd


def adjoint_workspace(solver, n_shots):
    ny, nx = solver.model.shape
    n = int(_native.lib().swfd_adjoint_workspace_floats(ny, nx, n_shots))
    return torch.empty(n, device=solver.model.device, dtype=_F32)


def adjoint_run(solver, sampler, n_shots, residual_nsr, step_hi, step_lo, *, state, work,
                hist=None, hist_first=0, hist_every=1, image_scale=0.0, grad=None,
                workspace=None, hist_lap=False, recon=None, recon_work=None, bnd=None,
                bnd_first=0, sources=None):
    """Adjoint steps ``step_hi-1 .. step_lo`` in one native call.

    ``state`` = (lam_old, lam) before the run; the state after the run is
    written back into ``state``'s tensors. ``work`` holds three or four
    scratch fields (may include the state tensors); with four the run takes
    several steps per launch. ``hist_lap``: the history holds Laplacians
    (``forward_run(hist_lap=True)``).

    Instead of ``hist``, ``recon`` = (field after step step_hi-1, field after
    step step_hi-2) of the forward run (read only) makes the launches rebuild
    the forward fields backwards in time and image against them; it needs
    ``recon_work`` (four fields), the forward run's boundary ring ``bnd`` and
    its ``sources`` = (rows_i32, cols_i32, values, src_weights or None).
    """
    dev = solver.model.device
    solver._sync_medium()
    if (hist is not None or recon is not None) and workspace is None:
        workspace = adjoint_workspace(solver, n_shots)
    rk = {}
    if recon is not None:
        if recon_work is None or len(recon_work) != 4 or bnd is None or sources is None:
            raise ValueError("reconstruction needs four recon_work fields, the ring and the "
                             "sources.")
        rows, cols, values, src_w = sources
        rk = dict(recon_next=_native.ptr(recon[0]), recon_cur=_native.ptr(recon[1]),
                  recon_work0=_native.ptr(recon_work[0]), recon_work1=_native.ptr(recon_work[1]),
                  recon_work2=_native.ptr(recon_work[2]), recon_work3=_native.ptr(recon_work[3]),
                  bnd=_native.ptr(_check_ring(bnd, n_shots, solver.model.shape)),
                  bnd_first=bnd_first, bnd_slots=bnd.shape[0], src_rows=_native.ptr(rows),
                  src_cols=_native.ptr(cols), src_values=_native.ptr(values),
                  src_weights=_native.ptr(_check_src_weights(src_w, n_shots)))
    args = _native.AdjointArgs(
        n_shots=n_shots, step_hi=step_hi, step_lo=step_lo,
        init_old=_native.ptr(state[0]), init_u=_native.ptr(state[1]),
        work0=_native.ptr(work[0]), work1=_native.ptr(work[1]), work2=_native.ptr(work[2]),
        work3=_native.ptr(work[3]) if len(work) > 3 else None,
        final_old=_native.ptr(state[0]), final_u=_native.ptr(state[1]),
        residual=_native.ptr(residual_nsr), hist=_native.ptr(hist), hist_first=hist_first,
        hist_every=hist_every, hist_slots=0 if hist is None else hist.shape[0],
        image_scale=float(image_scale), grad=_native.ptr(grad),
        workspace=_native.ptr(workspace), hist_lap=1 if hist_lap else 0, **rk)
    _native.check(_native.lib().swfd_adjoint(solver._grid, solver._medium, sampler._rec, args,
                                             _native.stream_of(dev)))


def simulate_shot(solver: TorchFDTD2D, sampler: TorchRingSampler, pulse, source_row: int,
                  source_col: int, n_steps: int, snapshot_stride: int = 0, source_weights=None):
    """One shot; returns traces (R, N) on the device and host snapshots.

    ``source_weights`` (4,): a bilinear source whose footprint corner is
    (source_row, source_col) (``RingArray.bilinear_sources``)."""
    solver.reset()
    if source_weights is None:
        solver.set_source(source_row, source_col)
    else:
        solver.set_bilinear_source(source_row, source_col, source_weights)
    dev = solver.model.device
    n_steps = int(n_steps)
    if snapshot_stride:
        traces = torch.empty((sampler.n_elements, n_steps), device=dev, dtype=_F32)
        snapshots = []
        p = torch.as_tensor(pulse).reshape(-1).to(device=dev, dtype=_F32)
        for step in range(n_steps):
            field = solver.inject_and_step(p[step]) if step < p.numel() else solver.step()
            traces[:, step] = sampler.record(field)
            if step % snapshot_stride == 0:
                snapshots.append((field.detach().cpu().numpy().copy(), step))
        return traces, snapshots
    values = _source_values_tensor(pulse, n_steps, dev)
    rc = torch.tensor([[source_row], [source_col]], dtype=torch.int32).to(dev)
    samples = torch.empty((n_steps, 1, sampler.n_elements), device=dev, dtype=_F32)
    u_old, u = forward_run(solver, 1, rc[0], rc[1], values, n_steps, samples=samples,
                           sampler=sampler, src_weights=solver._src_w)
    solver.u_old = u_old[0].clone()
    solver.u = u[0].clone()
    return samples[:, 0, :].transpose(0, 1).contiguous(), []


class TorchFDTD2DBatch:
    """State of independent shots advanced together (torch_backend.py:601-733)."""

    def __init__(self, solver: TorchFDTD2D, n_shots: int):
        if n_shots < 1:
            raise ValueError("n_shots must be positive.")
        self.model = solver.model
        self._solver = solver
        self.dt2 = solver.dt2
        self.n_shots = int(n_shots)
        dev = self.model.device
        self.u_old = torch.zeros((self.n_shots, *self.model.shape), device=dev, dtype=_F32)
        self.u = torch.zeros_like(self.u_old)
        self.batch_indices = torch.arange(self.n_shots, device=dev)
        self.source_rows = torch.zeros(self.n_shots, device=dev, dtype=torch.long)
        self.source_cols = torch.zeros(self.n_shots, device=dev, dtype=torch.long)
        self._rows_i32 = torch.zeros(self.n_shots, device=dev, dtype=torch.int32)
        self._cols_i32 = torch.zeros(self.n_shots, device=dev, dtype=torch.int32)
        self.compiled = solver.compiled

    def reset(self):
        self.u_old = torch.zeros_like(self.u_old)
        self.u = torch.zeros_like(self.u)

    def set_sources(self, rows, cols):
        rows = torch.as_tensor(np.asarray(rows) if not isinstance(rows, torch.Tensor) else rows,
                               dtype=torch.long).to(self.model.device)
        cols = torch.as_tensor(np.asarray(cols) if not isinstance(cols, torch.Tensor) else cols,
                               dtype=torch.long).to(self.model.device)
        if tuple(rows.shape) != (self.n_shots,) or tuple(cols.shape) != (self.n_shots,):
            raise ValueError("One source row and column is required per batched shot.")
        self.source_rows.copy_(rows)
        self.source_cols.copy_(cols)
        self._rows_i32.copy_(rows)
        self._cols_i32.copy_(cols)

    def _new_step_args(self, adjoint):
        out = torch.empty_like(self.u)
        args = _native.StepArgs(n_shots=self.n_shots, adjoint=int(adjoint),
                                u_old=_native.ptr(self.u_old), u=_native.ptr(self.u),
                                u_next=_native.ptr(out))
        return args, out

    def _run(self, args, rec=None):
        solver = self._solver
        solver._sync_medium()
        _native.check(_native.lib().swfd_step(solver._grid, solver._medium, rec, args,
                                              _native.stream_of(self.model.device)))

    def step(self, source_value):
        args, out = self._new_step_args(False)
        vt = self._solver._value_tensor(source_value)
        args.src_rows = _native.ptr(self._rows_i32)
        args.src_cols = _native.ptr(self._cols_i32)
        args.src_values = _native.ptr(vt)
        self._run(args)
        self.u_old, self.u = self.u, out
        return self.u

    def step_record(self, source_value, sampler):
        """Advance every shot one step and return its receiver samples."""
        args, out = self._new_step_args(False)
        vt = self._solver._value_tensor(source_value)
        samples = torch.empty((self.n_shots, sampler.n_elements), device=self.model.device,
                              dtype=_F32)
        args.src_rows = _native.ptr(self._rows_i32)
        args.src_cols = _native.ptr(self._cols_i32)
        args.src_values = _native.ptr(vt)
        args.samples = _native.ptr(samples)
        self._run(args, sampler._rec)
        self.u_old, self.u = self.u, out
        return self.u, samples

    def _residual_tensor(self, residual, sampler):
        residual = torch.as_tensor(residual, device=self.model.device, dtype=_F32)
        if tuple(residual.shape) != (self.n_shots, sampler.n_elements):
            raise ValueError("Batched residuals must have shape (shots, receivers).")
        return residual.contiguous()

    def adjoint_step_residual(self, residual, sampler):
        """Adjoint step driven by ``(shots, receivers)`` residual samples."""
        args, out = self._new_step_args(True)
        residual = self._residual_tensor(residual, sampler)
        args.residual = _native.ptr(residual)
        self._run(args, sampler._rec)
        self.u_old, self.u = self.u, out
        return self.u

    def adjoint_step_image(self, residual, sampler, forward_field, gradient, scale):
        """Adjoint step plus ``scale * sum(adjoint * L[forward_field])``; returns a new gradient."""
        args, out = self._new_step_args(True)
        residual = self._residual_tensor(residual, sampler)
        fwd = torch.as_tensor(forward_field, device=self.model.device).contiguous()
        if tuple(fwd.shape) != tuple(self.u.shape):
            raise ValueError("forward_field must have shape (shots, ny, nx).")
        if fwd.dtype != _F32:  # a reduced-precision history (u_fwd.to(dtype) in the reference)
            fwd = convert_history(fwd, torch.empty(fwd.shape, device=fwd.device, dtype=_F32))
        g_in = torch.as_tensor(gradient, device=self.model.device, dtype=_F32).contiguous()
        g_out = torch.empty_like(g_in)
        args.residual = _native.ptr(residual)
        args.image_field = _native.ptr(fwd)
        args.grad_in = _native.ptr(g_in)
        args.grad_out = _native.ptr(g_out)
        args.image_scale = float(scale)
        self._run(args, sampler._rec)
        self.u_old, self.u = self.u, out
        return g_out

    def adjoint_step(self, source_fields):
        args, out = self._new_step_args(True)
        dense = torch.as_tensor(source_fields, device=self.model.device, dtype=_F32).contiguous()
        if tuple(dense.shape) != tuple(self.u.shape):
            raise ValueError("Dense source fields must have shape (shots, ny, nx).")
        args.dense = _native.ptr(dense)
        self._run(args)
        self.u_old, self.u = self.u, out
        return self.u


def simulate_shots_batch(solver: TorchFDTD2D, sampler: TorchRingSampler, pulse, source_rows,
                         source_cols, n_steps: int, batched: TorchFDTD2DBatch | None = None,
                         source_values=None, source_weights=None):
    """Independent shots with a shared waveform; returns traces (S, R, N) and the final u.

    ``source_weights`` (S, 4): bilinear sources whose footprint corners are the
    rows and columns given (``RingArray.bilinear_sources``)."""
    if batched is None:
        batched = TorchFDTD2DBatch(solver, len(source_rows))
    elif batched.n_shots != len(source_rows):
        raise ValueError("Batched solver size does not match the source batch.")
    batched.set_sources(source_rows, source_cols)
    n_steps = int(n_steps)
    dev = solver.model.device
    values = _source_values_tensor(pulse, n_steps, dev, source_values)
    samples = torch.empty((n_steps, batched.n_shots, sampler.n_elements), device=dev, dtype=_F32)
    u_old, u = forward_run(solver, batched.n_shots, batched._rows_i32, batched._cols_i32,
                           values, n_steps, samples=samples, sampler=sampler,
                           src_weights=_weights_tensor(source_weights, batched.n_shots, dev))
    batched.u_old, batched.u = u_old, u
    traces = samples.permute(1, 2, 0).contiguous()
    return traces, batched.u


def auto_batch_size(device, n_shots, bytes_per_shot, fallback=16, memory_fraction=0.75):
    """Largest even split of ``n_shots`` whose working set fits in device memory."""
    device = torch.device(device)
    if device.type != "cuda":
        limit = fallback
    else:
        free, _ = torch.cuda.mem_get_info(device)
        cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
        limit = int(memory_fraction * (free + cached) // max(int(bytes_per_shot), 1))
    limit = max(1, min(int(n_shots), limit))
    n_batches = -(-int(n_shots) // limit)
    return -(-int(n_shots) // n_batches)
