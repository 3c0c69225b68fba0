"""Full-waveform inversion on the ring-array FDTD model (host + device)."""

from .medium import c_from_m, clip_m, interior_mask, m_from_c
from .misfit import mute_transmitter, tikhonov, total_variation, waveform_misfit

__all__ = [
    "c_from_m",
    "clip_m",
    "interior_mask",
    "m_from_c",
    "mute_transmitter",
    "tikhonov",
    "total_variation",
    "waveform_misfit",
]
