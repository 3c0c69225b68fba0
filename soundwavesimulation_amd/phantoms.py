"""Seeded synthetic sound-speed phantoms and starting models.

The rotated-ellipse membership test matches the reference rasteriser
(``fdtd2d/phantoms.py:27-35``). The random-ellipse generator itself has no
reference counterpart: water background, ``k`` ellipses with centres uniform
in a disk of radius ``centre_fraction * ring_radius``, semi-axes uniform in
``semi_axes_m``, rotation uniform in [0, pi) and speed uniform in
``speed_range``; later ellipses overwrite earlier ones.

``blurred_start`` restates the reference blurred-truth start
(``invert_ring.py:1179-1189``): Gaussian blur of slowness squared with
``mode="nearest"``, exterior reset to the background.
"""

from __future__ import annotations

import numpy as np

from .experiment import axis_centers


def ellipse_mask(x, y, x0, y0, a, b, phi_rad):
    """Points of ``(x, y)`` inside the ellipse (semi-axes a, b, rotation phi)."""
    dx = x - x0
    dy = y - y0
    cos_p = np.cos(phi_rad)
    sin_p = np.sin(phi_rad)
    xr = dx * cos_p + dy * sin_p
    yr = -dx * sin_p + dy * cos_p
    return (xr / a) ** 2 + (yr / b) ** 2 <= 1.0


def random_ellipses(n, seed, ring_radius_m, semi_axes_m=(5.0e-3, 30.0e-3),
                    speed_range=(1400.0, 1650.0), centre_fraction=0.8):
    """Seeded ellipse parameters ``(x0, y0, a, b, phi, c)``, shape (n, 6)."""
    rng = np.random.default_rng(seed)
    out = np.empty((int(n), 6), dtype=float)
    for i in range(int(n)):
        r = centre_fraction * ring_radius_m * np.sqrt(rng.uniform())
        t = rng.uniform(0.0, 2.0 * np.pi)
        out[i, 0] = r * np.cos(t)
        out[i, 1] = r * np.sin(t)
        out[i, 2] = rng.uniform(*semi_axes_m)
        out[i, 3] = rng.uniform(*semi_axes_m)
        out[i, 4] = rng.uniform(0.0, np.pi)
        out[i, 5] = rng.uniform(*speed_range)
    return out


def ellipse_phantom(nx, ny, spacing_m, n_ellipses, seed, ring_radius_m,
                    c_background=1500.0, **kwargs):
    """Speed map (ny, nx) float64 of seeded ellipses in a water background."""
    x_m = axis_centers(nx, spacing_m)
    y_m = axis_centers(ny, spacing_m)
    x, y = np.meshgrid(x_m, y_m, indexing="xy")
    c = np.full((ny, nx), float(c_background))
    for x0, y0, a, b, phi, speed in random_ellipses(n_ellipses, seed, ring_radius_m, **kwargs):
        c[ellipse_mask(x, y, x0, y0, a, b, phi)] = speed
    return c


def blurred_start(c_true, sigma_m, spacing_m, mask, c_background=1500.0):
    """Slowness squared of ``c_true`` blurred by a Gaussian of ``sigma_m``."""
    from scipy.ndimage import gaussian_filter

    m = 1.0 / (np.asarray(c_true, dtype=float) ** 2)
    m0 = gaussian_filter(m, float(sigma_m) / float(spacing_m), mode="nearest")
    m0[~np.asarray(mask, dtype=bool)] = 1.0 / float(c_background) ** 2
    return m0
