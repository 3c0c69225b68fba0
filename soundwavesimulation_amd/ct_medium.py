"""CT-derived speed maps placed on a square ring-FDTD grid (host side).

Mirrors ``fdtd2d/ct_medium.py`` of the reference: the ``.npz`` medium format
(``speed_m_s``, ``body_mask``, ``spacing_mm``; validated as in
``ct_medium.py:26-46``) and ``load_ct_ring_grid`` (``ct_medium.py:100-208``),
which centres the body's bounding box on the grid, picks an isotropic output
spacing no finer than the CT pixels and coarse enough for the body plus a
clearance to fit inside a ring that keeps ``edge_margin_pixels`` from the grid
edge, samples speed and body with nearest-neighbour lookups, and pads outside
the body with a coupling speed. The result feeds the same ``RingArray`` /
``TorchScalarWave2D`` path as the synthetic phantoms.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .experiment import axis_centers

REQUIRED_KEYS = ("body_mask", "spacing_mm", "speed_m_s")


@dataclass(frozen=True)
class CTRingGrid:
    """A centred CT medium and its ring radius on a square grid."""

    c: np.ndarray
    body_mask: np.ndarray
    x_m: np.ndarray
    y_m: np.ndarray
    spacing_m: float
    ring_radius_m: float
    source_shape: tuple
    source_spacing_mm: tuple


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


def body_extent(body, spacing_mm):
    """Centre (row, col) of the body's bounding box, in source pixels, and the
    largest physical distance (mm) from it to a body pixel."""
    rows, cols = np.nonzero(body)
    r0 = 0.5 * (rows.min() + rows.max())
    c0 = 0.5 * (cols.min() + cols.max())
    dy, dx = spacing_mm
    radius = np.hypot((rows - r0) * dy, (cols - c0) * dx).max()
    return float(r0), float(c0), float(radius)


def _resample(values, shape, spacing_mm, centre, source_spacing_mm, fill, order=0):
    """Values at the output grid's pixel centres, the grid centred on
    ``centre`` (source pixel coordinates); nearest neighbour, ``fill`` outside."""
    from scipy.ndimage import map_coordinates

    ny, nx = shape
    # (offset * spacing) / source spacing, in that order: nearest-neighbour
    # picks at half-pixel ties then fall as the reference's do
    rows = centre[0] + (np.arange(ny) - 0.5 * (ny - 1)) * spacing_mm / source_spacing_mm[0]
    cols = centre[1] + (np.arange(nx) - 0.5 * (nx - 1)) * spacing_mm / source_spacing_mm[1]
    rr, cc = np.meshgrid(rows, cols, indexing="ij")
    return map_coordinates(values, (rr, cc), order=order, mode="constant", cval=fill,
                           prefilter=False)


def _place(speed, body, shape, spacing_mm, centre, source_spacing_mm, padding):
    out_body = _resample(body.astype(np.uint8), shape, spacing_mm, centre, source_spacing_mm,
                         0).astype(bool)
    out_speed = _resample(speed, shape, spacing_mm, centre, source_spacing_mm, padding)
    out_speed[~out_body] = padding
    rows, cols = np.nonzero(out_body)
    mid = 0.5 * (shape[1] - 1)
    radius_px = float(np.hypot(rows - mid, cols - mid).max())
    return out_speed, out_body, radius_px


def load_ct_ring_grid(path, grid_shape=(512, 512), padding_speed_m_s=1480.0,
                      ring_clearance_mm=10.0, edge_margin_pixels=8):
    """Load a CT medium and place it, centred and downsampled, on a square grid
    inside a ring with ``ring_clearance_mm`` of clearance around the body."""
    speed, body, src_spacing = load_ct_npz(Path(path))
    ny, nx = (int(v) for v in grid_shape)
    if ny != nx or ny < 16:
        raise ValueError("The CT FDTD grid must be square and at least 16x16.")
    if not np.isfinite(padding_speed_m_s) or padding_speed_m_s <= 0:
        raise ValueError("padding_speed_m_s must be positive.")
    if ring_clearance_mm <= 0:
        raise ValueError("ring_clearance_mm must be positive.")
    if edge_margin_pixels < 3 or edge_margin_pixels >= nx // 4:
        raise ValueError("edge_margin_pixels must be at least 3 and below N/4.")
    padding = float(padding_speed_m_s)
    clearance = float(ring_clearance_mm)

    r0, c0, body_radius_mm = body_extent(body, src_spacing)
    ring_px = 0.5 * (nx - 1) - float(edge_margin_pixels)
    spacing_mm = max(src_spacing[0], src_spacing[1], (body_radius_mm + clearance) / ring_px)
    out_speed, out_body, radius_px = _place(speed, body, (ny, nx), spacing_mm, (r0, c0),
                                            src_spacing, padding)
    ring_mm = radius_px * spacing_mm + clearance
    limit_mm = ring_px * spacing_mm
    if ring_mm > limit_mm:
        # nearest-neighbour sampling can widen the body by part of a pixel:
        # coarsen the spacing by the overshoot and sample again
        spacing_mm *= ring_mm / limit_mm
        out_speed, out_body, radius_px = _place(speed, body, (ny, nx), spacing_mm, (r0, c0),
                                                src_spacing, padding)
        ring_mm = radius_px * spacing_mm + clearance

    spacing_m = 1.0e-3 * spacing_mm
    return CTRingGrid(c=out_speed, body_mask=out_body, x_m=axis_centers(nx, spacing_m),
                      y_m=axis_centers(ny, spacing_m), spacing_m=spacing_m,
                      ring_radius_m=1.0e-3 * ring_mm, source_shape=tuple(int(v) for v in
                                                                         speed.shape),
                      source_spacing_mm=src_spacing)
