"""Source waveforms sampled on the FDTD time grid.

``linear_chirp`` restates the reference Hann-windowed linear FM pulse
(``fdtd2d/sources.py:6-26``): ``n = max(round(T/dt), 2)`` samples at
``t = k dt``, sweep rate ``(f1 - f0) / ((n - 1) dt)``.

``ricker`` is the zero-phase Ricker wavelet
``(1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2)``, ``tau = t - delay`` with the
default delay ``1.5 / f``. The reference has no Ricker; it is passed to the
solvers through the same per-step ``source_values`` slot as the chirp.
"""

from __future__ import annotations

import numpy as np


def hann_window(n):
    if n <= 1:
        return np.ones(max(n, 0), dtype=float)
    return 0.5 - 0.5 * np.cos(2.0 * np.pi * np.arange(n) / (n - 1))


def linear_chirp(dt, duration_s, f_start_hz, f_end_hz):
    """Hann-windowed linear chirp starting at t = 0 (float64 NumPy)."""
    if duration_s <= 0:
        raise ValueError("Chirp duration must be positive.")
    if f_start_hz <= 0 or f_end_hz <= 0:
        raise ValueError("Chirp frequencies must be positive.")
    n_pulse = max(int(round(duration_s / dt)), 2)
    t = np.arange(n_pulse) * dt
    sweep = (f_end_hz - f_start_hz) / t[-1]
    phase = 2.0 * np.pi * (f_start_hz * t + 0.5 * sweep * t**2)
    return hann_window(n_pulse) * np.sin(phase)


def ricker(dt, f_peak_hz, n_samples=None, delay_s=None):
    """Ricker wavelet samples at ``t = k dt`` (float64 NumPy).

    By default it covers ``[0, 2 * delay]`` with ``delay = 1.5 / f``, i.e. the
    whole support where the wavelet exceeds about 1e-9 of its peak.
    """
    if f_peak_hz <= 0.0:
        raise ValueError("Ricker peak frequency must be positive.")
    if dt <= 0.0:
        raise ValueError("dt must be positive.")
    delay = 1.5 / f_peak_hz if delay_s is None else float(delay_s)
    if n_samples is None:
        n_samples = int(np.ceil(2.0 * delay / dt)) + 1
    t = np.arange(int(n_samples)) * dt
    arg = (np.pi * f_peak_hz * (t - delay)) ** 2
    return (1.0 - 2.0 * arg) * np.exp(-arg)


def source_values(pulse, n_steps):
    """Zero-padded (or truncated) per-step source samples of length ``n_steps``."""
    pulse = np.asarray(pulse, dtype=float).ravel()
    out = np.zeros(int(n_steps), dtype=float)
    k = min(pulse.size, out.size)
    out[:k] = pulse[:k]
    return out
