"""Slowness-squared parameterisation and the interior inversion mask.

Restates the reference ``fdtd2d/inversion/medium.py:15-40``: ``m = 1/c^2``,
``c = 1/sqrt(m)``, clipping to ``[1/c_max^2, 1/c_min^2]`` and the interior
mask ``r < radius - margin``.
"""

from __future__ import annotations

import numpy as np


def m_from_c(c):
    c = np.asarray(c, dtype=float)
    return 1.0 / (c * c)


def c_from_m(m):
    m = np.asarray(m, dtype=float)
    return 1.0 / np.sqrt(m)


def clip_m(m, c_min, c_max):
    m_min = 1.0 / (float(c_max) ** 2)
    m_max = 1.0 / (float(c_min) ** 2)
    return np.clip(m, m_min, m_max)


def interior_mask(x_m, y_m, radius_m, margin_m, center=(0.0, 0.0)):
    x, y = np.meshgrid(x_m, y_m, indexing="xy")
    ddx = x - float(center[0])
    ddy = y - float(center[1])
    radius = float(radius_m) - float(margin_m)
    if radius <= 0.0:
        raise ValueError("Interior radius must be positive; reduce the margin.")
    return ddx * ddx + ddy * ddy < radius * radius
