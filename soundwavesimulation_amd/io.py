"""On-disk formats shared with the reference (``.npz`` archives).

* ``save_traces`` / ``load_traces``: a forward simulation's receiver traces and
  acquisition, with the keys of the reference's ``save_traces``
  (``2DRingFDTD.py:291-326``): traces, time_s, pulse, dt, tx, theta, x, y,
  radius_m, sound_speed_m_s, torch_device, torch_dtype, torch_compiled,
  transmitter_indices.
* ``save_model`` / ``load_model``: an inversion result with the keys of the
  reference's ``save_model`` (``invert_ring.py:1262-1270``, compressed):
  c_true, c_est, mask, x_m, y_m, element_x, element_y, ring_radius_m,
  ring_center, misfit, optimizer -- the input of its ``--replot`` path.

Archives are plain arrays; ``load_*`` read them with ``allow_pickle=False``.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np


def save_traces(path, traces, time_s, array, tx, pulse, dt, c_map=None, device="cpu",
                dtype="float32", compiled=False, transmitter_indices=None):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez(
        path, traces=np.asarray(traces), time_s=np.asarray(time_s), pulse=np.asarray(pulse),
        dt=dt, tx=tx, theta=array.theta, x=array.x, y=array.y, radius_m=array.radius_m,
        sound_speed_m_s=c_map, torch_device=np.asarray(str(device)),
        torch_dtype=np.asarray(str(dtype)), torch_compiled=np.asarray(bool(compiled)),
        transmitter_indices=(np.asarray(transmitter_indices, dtype=int)
                             if transmitter_indices is not None
                             else np.asarray([tx], dtype=int)),
    )
    return path


def load_traces(path):
    """Every array of a traces archive. An entry stored as a Python object (the
    reference writes ``sound_speed_m_s=None`` when no map is given) would need
    unpickling, so it is left out rather than loaded."""
    out = {}
    with np.load(path, allow_pickle=False) as f:
        for k in f.files:
            try:
                out[k] = f[k]
            except ValueError:  # object array: refused without pickle
                continue
    return out


def save_model(path, c_true, c_est, history, x_m, y_m, array, mask, optimizer):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(
        path, c_true=c_true, c_est=c_est, mask=mask, x_m=x_m, y_m=y_m,
        element_x=array.x, element_y=array.y, ring_radius_m=array.radius_m,
        ring_center=np.asarray(array.center), misfit=np.asarray(history["misfit"]),
        optimizer=np.asarray(optimizer),
    )
    return path


def load_model(path):
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}
