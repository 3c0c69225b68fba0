"""Waveform misfit and the Tikhonov / total-variation regularisers on m.

Host NumPy, float64, with the definitions of ``fdtd2d/inversion/misfit.py:6-71``:

* waveform misfit ``1/2 ||P u - d||^2`` over every receiver channel but the
  transmitter's own;
* Tikhonov ``1/2 alpha ||D m||^2`` and smoothed isotropic total variation
  ``alpha sum_cells (sqrt(|D m|^2 + eps^2) - eps)``, where ``D`` takes the
  forward differences along x and y and keeps a difference only when both of
  its cells lie in the mask (so the fixed exterior adds no jump penalty).

Both regularisers are written with ``D`` and its transpose: the gradient of
``1/2 alpha ||D m||^2`` is ``alpha D^T D m``, and that of the TV term is
``D^T (alpha D m / sqrt(|D m|^2 + eps^2))``; cells outside the mask get zero.
"""

from __future__ import annotations

import numpy as np


def _kept_channels(n_channels, tx):
    keep = np.ones(n_channels, dtype=bool)
    keep[tx] = False
    return keep


def mute_transmitter(traces, tx):
    """Float64 copy of ``traces`` (channels first) with channel ``tx`` zeroed."""
    traces = np.asarray(traces, dtype=float)
    keep = _kept_channels(traces.shape[0], tx)
    return np.where(keep.reshape((-1,) + (1,) * (traces.ndim - 1)), traces, 0.0)


def waveform_misfit(predicted, observed, tx):
    """``1/2 sum (predicted - observed)^2`` over every channel but ``tx``."""
    diff = np.asarray(predicted, dtype=float) - np.asarray(observed, dtype=float)
    kept = diff[_kept_channels(diff.shape[0], tx)]
    return 0.5 * float(np.vdot(kept, kept))


class MaskedDifferences:
    """Forward differences of a 2D field kept where both cells lie in ``mask``.

    ``apply(m)`` returns ``(dx, dy)`` shaped like ``m`` (zero in the last column /
    row and wherever a cell pair leaves the mask); ``transpose(dx, dy)`` is
    the adjoint map back to a field.
    """

    def __init__(self, mask):
        mask = np.asarray(mask, dtype=bool)
        self.keep_x = np.zeros(mask.shape, dtype=bool)
        self.keep_y = np.zeros(mask.shape, dtype=bool)
        self.keep_x[:, :-1] = mask[:, :-1] & mask[:, 1:]
        self.keep_y[:-1, :] = mask[:-1, :] & mask[1:, :]

    def apply(self, m):
        dx = np.zeros(m.shape)
        dy = np.zeros(m.shape)
        dx[:, :-1] = m[:, 1:] - m[:, :-1]
        dy[:-1, :] = m[1:, :] - m[:-1, :]
        return np.where(self.keep_x, dx, 0.0), np.where(self.keep_y, dy, 0.0)

    def transpose(self, dx, dy):
        dx = np.where(self.keep_x, dx, 0.0)
        dy = np.where(self.keep_y, dy, 0.0)
        out = -(dx + dy)
        out[:, 1:] += dx[:, :-1]
        out[1:, :] += dy[:-1, :]
        return out


def tikhonov(m, mask, alpha):
    """Value and gradient of ``1/2 alpha ||D m||^2`` on the mask."""
    m = np.asarray(m, dtype=float)
    mask = np.asarray(mask, dtype=bool)
    if alpha <= 0.0:
        return 0.0, np.zeros_like(m)
    D = MaskedDifferences(mask)
    dx, dy = D.apply(m)
    value = 0.5 * alpha * (float(np.vdot(dx, dx)) + float(np.vdot(dy, dy)))
    gradient = D.transpose(alpha * dx, alpha * dy)
    gradient[~mask] = 0.0
    return value, gradient


def total_variation(m, mask, alpha, eps):
    """Value and gradient of the smoothed isotropic TV on the mask."""
    m = np.asarray(m, dtype=float)
    mask = np.asarray(mask, dtype=bool)
    if alpha <= 0.0:
        return 0.0, np.zeros_like(m)
    D = MaskedDifferences(mask)
    dx, dy = D.apply(m)
    magnitude = np.sqrt(dx * dx + dy * dy + eps * eps)
    value = alpha * float(np.sum(magnitude - eps))
    gradient = D.transpose(alpha * dx / magnitude, alpha * dy / magnitude)
    gradient[~mask] = 0.0
    return value, gradient
