"""Learned surrogate: a spectral neural operator that maps the first FWI
gradient image to a starting speed model (see ``operator.py``), trained on
pairs the FDTD generates (``data.py``, ``train.py``)."""

from .operator import SpectralOperator, dft_matrices, init_params, mode_ky

__all__ = ["SpectralOperator", "dft_matrices", "init_params", "mode_ky"]
