"""The CT medium ``.npz`` format of the reference (host side).

``load_ct_npz`` reads and validates an archive with the keys and checks of
``fdtd2d/ct_medium.py:26-46`` (``speed_m_s``, ``body_mask``, ``spacing_mm``;
other keys such as ``hu`` or ``labels`` are ignored); ``save_ct_npz`` writes
one. Placing such a medium on a ring grid (the reference's CT embedding) is
outside the ported path.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

REQUIRED_KEYS = ("body_mask", "spacing_mm", "speed_m_s")


def load_ct_npz(path):
    """Read and validate a CT medium archive -> (speed, body, (dy_mm, dx_mm))."""
    with np.load(path, allow_pickle=False) as archive:
        missing = [k for k in REQUIRED_KEYS if k not in archive.files]
        if missing:
            raise ValueError(f"{path} is missing {', '.join(missing)}. "
                             "Use an .npz produced by ct_to_speed.py.")
        speed = np.array(archive["speed_m_s"], dtype=float)
        body = np.array(archive["body_mask"], dtype=bool)
        spacing = np.array(archive["spacing_mm"], dtype=float).reshape(-1)
    if speed.ndim != 2 or body.shape != speed.shape:
        raise ValueError("speed_m_s and body_mask must be matching 2D arrays.")
    if spacing.size != 2 or not np.isfinite(spacing).all() or (spacing <= 0).any():
        raise ValueError("spacing_mm must contain positive row and column spacing.")
    if not np.isfinite(speed).all() or (speed <= 0).any():
        raise ValueError("speed_m_s must contain finite positive values.")
    if not body.any():
        raise ValueError("body_mask contains no foreground pixels.")
    return speed, body, (float(spacing[0]), float(spacing[1]))


def save_ct_npz(path, speed_m_s, body_mask, spacing_mm):
    """Write a CT medium archive in the format ``load_ct_npz`` reads."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(path, speed_m_s=np.asarray(speed_m_s, dtype=float),
                        body_mask=np.asarray(body_mask, dtype=bool),
                        spacing_mm=np.asarray(spacing_mm, dtype=float))
    return path
