"""MI355X-native 2D ring-array FDTD and adjoint-state FWI.

Host-side geometry and waveform helpers are importable everywhere; the device
path lives in :mod:`soundwavesimulation_amd.torch_backend` (HIP kernels behind
a C-ABI library) and :mod:`soundwavesimulation_amd.inversion`.
"""

from .array import ArcArray, RingArray
from .experiment import axis_centers, cfl_number, n_steps_for_crossing, stable_dt
from .phantoms import blurred_start, ellipse_phantom
from .sources import linear_chirp, ricker, source_values

__all__ = [
    "ArcArray",
    "RingArray",
    "axis_centers",
    "blurred_start",
    "cfl_number",
    "ellipse_phantom",
    "linear_chirp",
    "n_steps_for_crossing",
    "ricker",
    "source_values",
    "stable_dt",
]
