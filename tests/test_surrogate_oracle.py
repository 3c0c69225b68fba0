"""CPU checks of the float64 operator restatement (oracle/surrogate_oracle.py).

What pins it: the spectral convolution reproduces single retained Fourier
modes exactly and removes the others (closed form, independent of how the
transform is computed), its mode mixing matches an explicit complex product,
and the hand-written backward pass matches central finite differences of the
forward pass for every parameter.
"""

import numpy as np
import pytest

from oracle import surrogate_oracle as so
from soundwavesimulation_amd.surrogate.operator import dft_matrices, init_params, mode_ky


def _identity_mixing(m, C):
    re = np.zeros((m * m, C, C))
    re[:, np.arange(C), np.arange(C)] = 1.0
    return re, np.zeros_like(re)


@pytest.mark.parametrize("ky,kx,keep", [(0, 0, True), (3, 2, True), (-4, 5, True), (2, 7, True),
                                        (-3, 0, True), (1, 8, False), (6, 1, False)])
def test_spectral_keeps_retained_modes_and_drops_the_rest(ky, kx, keep):
    n, m, C = 16, 8, 2
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    th = 2 * np.pi * (ky * y + kx * x) / n
    v = np.zeros((1, n, n, C))
    v[0, :, :, 0] = np.cos(th)
    v[0, :, :, 1] = np.sin(th)
    re, im = _identity_mixing(m, C)
    out, _ = so.spectral(v, re, im, m)
    expect = v if keep else np.zeros_like(v)
    if kx == 0 and ky == 0:
        expect = v.copy()
        expect[0, :, :, 1] = 0.0  # sin of a zero phase is zero anyway
    np.testing.assert_allclose(out, expect, atol=1e-12)


def test_spectral_mixing_is_a_complex_product_per_mode():
    n, m, C = 16, 4, 3
    rng = np.random.default_rng(0)
    v = rng.standard_normal((2, n, n, C))
    re, im = rng.standard_normal((2, m * m, C, C))
    out, X = so.spectral(v, re, im, m)
    # explicit: spectrum by matrix DFT (float64), per-mode complex mixing
    Fx, Fy, Gy, Gx = dft_matrices(n, m)
    ky = mode_ky(m)
    ex = np.exp(-2j * np.pi * np.outer(np.arange(m), np.arange(n)) / n)
    ey = np.exp(-2j * np.pi * np.outer(ky, np.arange(n)) / n)
    Xe = np.einsum("qy,jx,byxc->bqjc", ey, ex, v)
    np.testing.assert_allclose(X, Xe, atol=1e-10)
    R = (re + 1j * im).reshape(m, m, C, C)
    Y = np.einsum("bqji,qjio->bqjo", Xe, R)
    w = np.where(np.arange(m) == 0, 1.0, 2.0)
    ey_inv = np.exp(2j * np.pi * np.outer(np.arange(n), ky) / n)
    ex_inv = np.exp(2j * np.pi * np.outer(np.arange(n), np.arange(m)) / n) * w
    oe = np.real(np.einsum("yq,xj,bqjo->byxo", ey_inv, ex_inv, Y)) / (n * n)
    np.testing.assert_allclose(out, oe, atol=1e-10)
    # the cos / sin matrices of the device formulation give the same spectrum
    T1 = np.einsum("jx,byxc->byjc", Fx, v)
    P = np.einsum("ry,byjc->brjc", Fy, T1)
    Xre = P[:, :m, :m] + P[:, m:, m:]
    Xim = P[:, :m, m:] - P[:, m:, :m]
    np.testing.assert_allclose(Xre + 1j * Xim, Xe, atol=1e-10)


def test_backward_matches_finite_differences():
    n, m, C, H, L, B = 16, 4, 4, 8, 2, 2
    rng = np.random.default_rng(1)
    params = init_params(seed=3, width=C, modes=m, layers=L, hidden=H)
    for l in range(L):  # make the spectral path matter
        params[f"R{l}_re"] = rng.standard_normal(params[f"R{l}_re"].shape) * 0.3
        params[f"R{l}_im"] = rng.standard_normal(params[f"R{l}_im"].shape) * 0.3
        params[f"b{l}"] = rng.standard_normal(C) * 0.1
    params["bq1"] = rng.standard_normal(H) * 0.1
    inp = rng.standard_normal((B, n, n, 3))
    dout = rng.standard_normal((B, n, n))
    out, saved = so.forward(params, inp, modes=m, layers=L)
    grads = so.backward(params, saved, dout, modes=m, layers=L)
    assert set(grads) == set(params)

    def objective(p):
        return float((so.forward(p, inp, modes=m, layers=L)[0] * dout).sum())

    eps = 1e-6
    for name, val in params.items():
        direction = rng.standard_normal(val.shape)
        plus = dict(params)
        minus = dict(params)
        plus[name] = val + eps * direction
        minus[name] = val - eps * direction
        fd = (objective(plus) - objective(minus)) / (2 * eps)
        an = float((grads[name] * direction).sum())
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, fd, an)


def test_mse_gradient():
    rng = np.random.default_rng(2)
    p, t = rng.standard_normal((2, 3, 5, 5))
    mask = (rng.random((3, 5, 5)) > 0.3).astype(float)
    loss, grad = so.mse(p, t, mask)
    eps = 1e-7
    e = np.zeros_like(p)
    e[1, 2, 3] = 1.0
    fd = (so.mse(p + eps * e, t, mask)[0] - so.mse(p - eps * e, t, mask)[0]) / (2 * eps)
    assert abs(fd - grad[1, 2, 3]) < 1e-6


def test_device_formulation_in_float64_matches_the_oracle():
    """The GEMM / repack sequence of the device operator, evaluated in float64."""
    n, m, C, B = 16, 4, 3, 2
    rng = np.random.default_rng(4)
    v = rng.standard_normal((B, n, n, C))
    re, im = rng.standard_normal((2, m * m, C, C))
    Fx, Fy, Gy, Gx = dft_matrices(n, m)
    T1 = np.einsum("jx,byxc->byjc", Fx, v)                    # (B, N, 2m, C)
    P = np.einsum("ry,byjc->brjc", Fy, T1)                    # (B, 2m, 2m, C)
    X = np.concatenate([P[:, :m, :m] + P[:, m:, m:], P[:, :m, m:] - P[:, m:, :m]], axis=-1)
    Rm = np.zeros((m * m, 2 * C, 2 * C))
    Rm[:, :C, :C], Rm[:, :C, C:], Rm[:, C:, :C], Rm[:, C:, C:] = re, im, -im, re
    Y = np.einsum("bmi,mio->bmo", X.reshape(B, m * m, 2 * C), Rm).reshape(B, m, m, 2 * C)
    U = np.zeros((B, 2 * m, 2 * m, C))
    U[:, :m, :m], U[:, :m, m:] = Y[..., :C], Y[..., C:]
    U[:, m:, :m], U[:, m:, m:] = Y[..., C:], -Y[..., :C]
    Z = np.einsum("yr,brjc->byjc", Gy, U)                     # (B, N, 2m, C)
    out = np.einsum("xj,byjc->byxc", Gx, Z)
    ref, _ = so.spectral(v, re, im, m)
    np.testing.assert_allclose(out, ref, atol=1e-10)
