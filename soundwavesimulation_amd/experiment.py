"""Grid and time-step helpers shared by every driver.

Closed forms follow the reference ``fdtd2d/experiment.py:6-18``:
cell-centred axes with the origin at the grid centre, the 2D second-order
CFL time step, and the ring-crossing step count.
"""

from __future__ import annotations

import math

import numpy as np


def axis_centers(n, spacing):
    """Cell-centre coordinates ``(i - (n - 1) / 2) * spacing`` for ``i < n``."""
    return (np.arange(n) - 0.5 * (n - 1)) * spacing


def stable_dt(c_max, dx, dy, cfl=0.45):
    """``dt = cfl / (c_max * sqrt(1/dx^2 + 1/dy^2))``."""
    return cfl / (c_max * np.sqrt(1.0 / dx**2 + 1.0 / dy**2))


def n_steps_for_crossing(radius_m, c_min, dt, extras=1.6):
    """Steps to cross the ring diameter at ``c_min`` times a tail margin."""
    return int(np.ceil(extras * (2.0 * radius_m / c_min) / dt))


def cfl_number(c_max, dt, dx, dy):
    """Courant number ``c_max * dt * sqrt(1/dx^2 + 1/dy^2)`` (must be < 1)."""
    return float(c_max * dt * math.sqrt(1.0 / dx**2 + 1.0 / dy**2))
