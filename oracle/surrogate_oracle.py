"""Float64 CPU restatement of the spectral neural operator (test infrastructure).

Only tests/ (and bench.py's CPU leg) use this module, as the checker of the
device operator in ``soundwavesimulation_amd/surrogate/operator.py``. It is
written independently of the device formulation: the truncated transforms go
through ``numpy.fft`` (the device uses explicit cos / sin matrices in GEMMs),
and the backward pass uses the adjoint relations of the FFT:

* ``K(v) = Re ifft2(S)`` with ``S[ky, kx] = w_kx Y[ky, kx]`` on the retained
  modes (``ky in [-m/2, m/2)``, ``kx in [0, m)``, ``w = 1`` at kx = 0 and 2
  otherwise) and ``Y[.., o] = sum_i fft2(v)[.., i] R[mode, i, o]``;
* backward: ``gY = w / N^2 fft2(g)`` on the retained modes,
  ``gX = gY conj(R)^T``, ``gR = sum_b conj(X) gY``, ``dv = Re(N^2 ifft2(gX))``.

The operator has no counterpart in the reference's code (its design
document, TFNO_Full_Ring_Plan.tex:200-209, describes a Fourier neural
operator); parity of this restatement is pinned by the closed form on single
Fourier modes and by finite differences (tests/test_surrogate_oracle.py).
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def _modes(n, m):
    ky = np.array([q if q < m // 2 else q - m for q in range(m)])
    return np.mod(ky, n), np.arange(m), np.where(np.arange(m) == 0, 1.0, 2.0)


def spectral(v, re, im, m):
    """K(v) for v (B, N, N, C); returns (out, X) with X the retained spectrum."""
    B, n, _, C = v.shape
    iy, ix, w = _modes(n, m)
    F = np.fft.fft2(v, axes=(1, 2))
    X = F[:, iy][:, :, ix]                                  # (B, m, m, C)
    R = (re + 1j * im).reshape(m, m, C, C)                  # mode (q, j), i, o
    Y = np.einsum("bqji,qjio->bqjo", X, R)
    S = np.zeros((B, n, n, C), dtype=complex)
    S[:, iy[:, None], ix[None, :]] = Y * w[None, None, :, None]
    return np.real(np.fft.ifft2(S, axes=(1, 2))), X


def spectral_backward(g, X, re, im, m):
    """Gradients (dv, d_re, d_im) of <g, K(v)> for the spectrum X of v."""
    B, n, _, C = g.shape
    iy, ix, w = _modes(n, m)
    G = np.fft.fft2(g, axes=(1, 2))
    gY = G[:, iy][:, :, ix] * (w[None, None, :, None] / (n * n))
    R = (re + 1j * im).reshape(m, m, C, C)
    gX = np.einsum("bqjo,qjio->bqji", gY, np.conj(R))
    gR = np.einsum("bqji,bqjo->qjio", np.conj(X), gY).reshape(m * m, C, C)
    S = np.zeros((B, n, n, C), dtype=complex)
    S[:, iy[:, None], ix[None, :]] = gX
    dv = np.real(np.fft.ifft2(S, axes=(1, 2))) * (n * n)
    return dv, np.real(gR), np.imag(gR)


def forward(params, inp, modes=16, layers=4):
    """Operator output (B, N, N) and the saved activations."""
    v = inp @ params["P"] + params["bp"]
    saved = {"inp": inp, "v": [v], "pre": [], "X": []}
    for l in range(layers):
        k, X = spectral(v, params[f"R{l}_re"], params[f"R{l}_im"], modes)
        pre = k + v @ params[f"W{l}"] + params[f"b{l}"]
        v = gelu(pre) if l < layers - 1 else pre
        saved["pre"].append(pre)
        saved["X"].append(X)
        saved["v"].append(v)
    hpre = v @ params["Q1"] + params["bq1"]
    h = gelu(hpre)
    out = (h @ params["Q2"] + params["bq2"])[..., 0]
    saved.update(hpre=hpre, h=h)
    return out, saved


def backward(params, saved, dout, modes=16, layers=4):
    """Gradients of sum(dout * out) for every parameter (dict of arrays)."""
    g = {}
    d = dout[..., None]                                      # (B, N, N, 1)
    h = saved["h"]
    g["Q2"] = np.einsum("bxyh,bxyo->ho", h, d)
    g["bq2"] = d.sum(axis=(0, 1, 2))
    dh = (d @ params["Q2"].T) * gelu_grad(saved["hpre"])
    v = saved["v"][layers]
    g["Q1"] = np.einsum("bxyc,bxyh->ch", v, dh)
    g["bq1"] = dh.sum(axis=(0, 1, 2))
    dv = dh @ params["Q1"].T
    for l in range(layers - 1, -1, -1):
        pre = saved["pre"][l]
        dpre = dv * gelu_grad(pre) if l < layers - 1 else dv
        v = saved["v"][l]
        g[f"b{l}"] = dpre.sum(axis=(0, 1, 2))
        g[f"W{l}"] = np.einsum("bxyi,bxyo->io", v, dpre)
        dvs, gre, gim = spectral_backward(dpre, saved["X"][l], params[f"R{l}_re"],
                                          params[f"R{l}_im"], modes)
        g[f"R{l}_re"], g[f"R{l}_im"] = gre, gim
        dv = dpre @ params[f"W{l}"].T + dvs
    inp = saved["inp"]
    g["P"] = np.einsum("bxyi,bxyc->ic", inp, dv)
    g["bp"] = dv.sum(axis=(0, 1, 2))
    return g


def mse(pred, target, mask=None):
    """Masked mean squared error and its gradient with respect to pred."""
    w = np.ones_like(pred) if mask is None else mask
    d = pred - target
    ws = w.sum()
    return float((w * d * d).sum() / ws), 2.0 * w * d / ws
