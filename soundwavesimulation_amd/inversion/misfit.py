"""Waveform misfit and the Tikhonov / total-variation regularisers on m.

Host NumPy, float64, restating ``fdtd2d/inversion/misfit.py:6-71``:
``1/2 ||P u - d||^2`` with the transmitter row muted; Tikhonov
``1/2 alpha ||grad m||^2`` and smoothed isotropic TV
``alpha sum(sqrt(|grad m|^2 + eps^2) - eps)``, both counting a forward
difference only when both of its pixels lie in the mask.
"""

from __future__ import annotations

import numpy as np


def mute_transmitter(traces, tx):
    muted = np.array(traces, dtype=float, copy=True)
    muted[tx] = 0.0
    return muted


def waveform_misfit(predicted, observed, tx):
    residual = mute_transmitter(np.asarray(predicted) - np.asarray(observed), tx)
    return 0.5 * float(np.sum(residual * residual))


def tikhonov(m, mask, alpha):
    """Value and gradient of ``1/2 alpha sum_(in-mask edges) (m_i - m_j)^2``."""
    m = np.asarray(m, dtype=float)
    mask = np.asarray(mask, dtype=bool)
    g = np.zeros_like(m)
    if alpha <= 0.0:
        return 0.0, g
    value = 0.0
    for axis in (1, 0):
        lo = [slice(None), slice(None)]
        hi = [slice(None), slice(None)]
        lo[axis] = slice(None, -1)
        hi[axis] = slice(1, None)
        lo, hi = tuple(lo), tuple(hi)
        both = mask[hi] & mask[lo]
        diff = m[hi] - m[lo]
        value += 0.5 * alpha * float(np.sum(diff[both] ** 2))
        contrib = np.where(both, alpha * diff, 0.0)
        g[hi] += contrib
        g[lo] -= contrib
    g[~mask] = 0.0
    return value, g


def total_variation(m, mask, alpha, eps):
    """Value and gradient of the smoothed isotropic TV on the mask."""
    m = np.asarray(m, dtype=float)
    mask = np.asarray(mask, dtype=bool)
    g = np.zeros_like(m)
    if alpha <= 0.0:
        return 0.0, g
    ddx = np.zeros_like(m)
    ddy = np.zeros_like(m)
    ddx[:, :-1] = np.where(mask[:, 1:] & mask[:, :-1], m[:, 1:] - m[:, :-1], 0.0)
    ddy[:-1, :] = np.where(mask[1:, :] & mask[:-1, :], m[1:, :] - m[:-1, :], 0.0)
    norm = np.sqrt(ddx * ddx + ddy * ddy + eps * eps)
    value = alpha * float(np.sum(norm - eps))
    gx = alpha * ddx / norm
    gy = alpha * ddy / norm
    g -= gx + gy
    g[:, 1:] += gx[:, :-1]
    g[1:, :] += gy[:-1, :]
    g[~mask] = 0.0
    return value, g
