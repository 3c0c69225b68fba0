"""Spectral neural operator on the device, built from the ``libswno`` kernels.

Architecture (the Fourier neural operator of the reference's design document,
``TFNO_Full_Ring_Plan.tex:200-209``, at the sizes of the surrogate workload:
width 32, 4 spectral layers, 16 x 16 retained modes, GELU):

* input ``(B, N, N, 3)``: the normalised first-gradient image and the x / y
  coordinates in [-1, 1]; lifting ``v0 = in P + bp`` to width C;
* layer l: ``pre = K_l(v_l) + v_l W_l + b_l``, ``v_{l+1} = GELU(pre)`` (no
  activation after the last layer), with the spectral convolution
  ``K(v) = Re IDFT_m(R . DFT_m(v))``: the 2-D DFT truncated to
  ``ky in [-m/2, m/2)``, ``kx in [0, m)``, a complex C x C matrix per mode,
  and the inverse over the same modes (weight 2 for ``kx > 0``, 1/N^2);
* projection ``h = GELU(v4 Q1 + bq1)`` (width 128), ``out = h Q2 + bq2``.

Every contraction is one call of the bf16-MFMA GEMM (fp32 accumulation):
the truncated DFT along x is a GEMM batched over (sample, row) with a
(2m x N) cos / -sin matrix, along y a GEMM batched over samples, the mode
mixing a GEMM batched over modes with the real 2C x 2C form of each complex
matrix, and the pointwise linears one GEMM over all pixels; the backward pass
uses the same kernel on transposed (strided) views, with split-K for the
weight gradients (reductions over every pixel of the batch). Activations are
channel-last ``(B, N, N, C)``: fp32, except, on the fused path, the lifting's
output v_0 and the layer outputs v_1 .. v_L, which are stored as bf16 (every
consumer rounds them to bf16 for an MFMA product, so the results are the same
bits with half the bytes; ``SWNO_BF16_ACT=0`` keeps them fp32).

``SpectralOperator.forward`` keeps what ``backward`` needs; ``backward``
returns the gradient of the loss for every parameter in one flat buffer laid
out like the parameters (so AdamW is one launch).
"""

from __future__ import annotations

import math
import os

import numpy as np
import torch

from . import _native

_F32 = torch.float32


def init_params(seed=0, width=32, modes=16, layers=4, hidden=128, in_ch=3):
    """Seeded initial parameters (float64 NumPy arrays, name -> array)."""
    rng = np.random.default_rng(seed)

    def dense(fan_in, fan_out):
        b = 1.0 / math.sqrt(fan_in)
        return rng.uniform(-b, b, (fan_in, fan_out))

    p = {"P": dense(in_ch, width), "bp": np.zeros(width)}
    scale = 1.0 / (width * width)
    for l in range(layers):
        p[f"R{l}_re"] = scale * rng.random((modes * modes, width, width))
        p[f"R{l}_im"] = scale * rng.random((modes * modes, width, width))
        p[f"W{l}"] = dense(width, width)
        p[f"b{l}"] = np.zeros(width)
    p["Q1"] = dense(width, hidden)
    p["bq1"] = np.zeros(hidden)
    p["Q2"] = dense(hidden, 1)
    p["bq2"] = np.zeros(1)
    return p


def mode_ky(modes):
    """Retained y wavenumbers, in the order of the mode axis: 0..m/2-1, -m/2..-1."""
    return np.array([q if q < modes // 2 else q - modes for q in range(modes)])


def dft_matrices(n, modes):
    """Truncated transform matrices (float64): Fx, Fy (2m x n), Gy, Gx (n x 2m)."""
    m = modes
    pos = np.arange(n)
    kx = np.arange(m)
    ky = mode_ky(m)
    ax = 2.0 * np.pi * np.outer(kx, pos) / n          # (m, n)
    ay = 2.0 * np.pi * np.outer(ky, pos) / n          # (m, n)
    Fx = np.concatenate([np.cos(ax), -np.sin(ax)])    # A_re, A_im rows
    Fy = np.concatenate([np.cos(ay), np.sin(ay)])
    Gy = np.concatenate([np.cos(ay), -np.sin(ay)]).T
    w = np.where(kx == 0, 1.0, 2.0)[:, None] / (n * n)
    Gx = np.concatenate([w * np.cos(ax), -w * np.sin(ax)]).T
    return Fx, Fy, Gy, Gx


def _ptr(t):
    return None if t is None else t.data_ptr()


def _is_bf16(t):
    return 1 if t.dtype == torch.bfloat16 else 0


class SpectralOperator:
    """Device parameters, buffers and forward / backward of the operator."""

    def __init__(self, params, n, device, *, width=32, modes=16, layers=4, hidden=128,
                 in_ch=3):
        if device.type != "cuda":
            raise ValueError("SpectralOperator runs on a HIP device only.")
        self.lib = _native.load_library()
        self.n, self.C, self.m, self.L, self.H, self.Cin = n, width, modes, layers, hidden, in_ch
        if n < 2 * modes:
            raise ValueError(f"grid {n} too small for {modes} modes per axis")
        self.device = device
        self.names = list(params.keys())
        sizes = [int(np.prod(params[k].shape)) for k in self.names]
        self.flat = torch.empty(sum(sizes), device=device, dtype=_F32)
        self.grad_flat = torch.zeros_like(self.flat)
        self.p, self.g = {}, {}
        off = 0
        for k, sz in zip(self.names, sizes):
            shape = params[k].shape
            self.p[k] = self.flat[off:off + sz].view(shape)
            self.g[k] = self.grad_flat[off:off + sz].view(shape)
            self.p[k].copy_(torch.as_tensor(np.asarray(params[k], dtype=np.float32)))
            off += sz
        Fx, Fy, Gy, Gx = dft_matrices(n, modes)
        dev = device
        self.Fx = torch.as_tensor(Fx, dtype=_F32, device=dev).contiguous()
        self.Fy = torch.as_tensor(Fy, dtype=_F32, device=dev).contiguous()
        self.Gy = torch.as_tensor(Gy, dtype=_F32, device=dev).contiguous()
        self.Gx = torch.as_tensor(Gx, dtype=_F32, device=dev).contiguous()
        self.rm = [torch.empty((modes * modes, 2 * width, 2 * width), device=dev, dtype=_F32)
                   for _ in range(layers)]
        self.drm = torch.empty_like(self.rm[0])
        # the projection head as one fused kernel each way (built for width 32, hidden 128)
        fuse = os.environ.get("SWNO_FUSED", "1") != "0"
        self.fused_head = fuse and (width, hidden) == (32, 128)
        # spectral layer tails (x inverse transform + pointwise linear + bias/GELU, and the
        # backward counterpart) as one kernel each way
        self.fused_layers = (fuse and width == 32 and n % 16 == 0 and n <= 128
                             and modes % 2 == 0 and modes <= 16)
        self._bufs = {}
        self.adam_m = torch.zeros_like(self.flat)
        self.adam_v = torch.zeros_like(self.flat)
        self.adam_step = 0

    # ---------------------------------------------------------------- helpers
    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _buffers(self, B):
        if B in self._bufs:
            return self._bufs[B]
        n, C, m, H, L = self.n, self.C, self.m, self.H, self.L
        rows = B * n * n
        e = lambda *s: torch.empty(s, device=self.device, dtype=_F32)  # noqa: E731
        # layer outputs in bf16 where only the fused kernels (which round them to bf16
        # for their products anyway) read them
        bf16_act = (self.fused_layers and self.fused_head
                    and os.environ.get("SWNO_BF16_ACT", "1") != "0")
        b = {
            "inp": e(rows, self.Cin),
            "v": [torch.empty((rows, C), device=self.device,
                              dtype=torch.bfloat16 if bf16_act else _F32)
                  for _ in range(L + 1)],
            "pre": [e(rows, C) for _ in range(L)],
            "hpre": e(rows, H), "h": e(rows, H), "out": e(rows, 1), "outpre": e(rows, 1),
            "T1": e(B * n * 2 * m * C), "P": e(B * 4 * m * m * C),
            "X": [e(B * m * m * 2 * C) for _ in range(L)], "Y": e(B * m * m * 2 * C),
            "U": e(B * 4 * m * m * C), "Z": e(B * n * 2 * m * C),
            "dv": e(rows, C), "dpre": e(rows, C), "dh": e(rows, H), "dout": e(rows, 1),
            "wgws": e(max(int(self.lib.swno_weight_grad_workspace_floats(rows, mm, nn))
                          for mm, nn in ((H, 1), (C, H), (C, C), (self.Cin, C)))),
            "headws": e(int(self.lib.swno_head_backward_workspace_floats(rows))),
            "msews": e(int(self.lib.swno_mse_workspace_floats(rows)) + 2),
            "loss": e(2),
        }
        self._bufs[B] = b
        return b

    def gemm(self, M, N, K, batch, A, a, Bm, b, Cm, c, alpha=1.0, beta=0.0, k_split=1, ws=None):
        """One GEMM call; ``a``, ``b``, ``c`` are (row, col, batch) element strides."""
        g = _native.GemmArgs(M, N, K, batch, k_split, _ptr(A), a[0], a[1], a[2], _ptr(Bm),
                             b[0], b[1], b[2], _ptr(Cm), c[0], c[1], c[2], alpha, beta)
        if k_split > 1:
            need = int(self.lib.swno_gemm_workspace_floats(g))
            if ws is None or ws.numel() < need:
                raise ValueError("GEMM workspace too small")
        _native.check(self.lib.swno_gemm(g, _ptr(ws), self._stream()))

    def weight_grad(self, rows, A, M, B, N, dW, db, ws):
        """dW (M x N) = A^T B and db = column sums of B over ``rows`` rows."""
        _native.check(self.lib.swno_weight_grad(rows, M, N, _ptr(A), M, _ptr(B), N, _ptr(dW),
                                                _ptr(db), _ptr(ws), _is_bf16(A),
                                                self._stream()))

    # ------------------------------------------------------------ spectral
    def _spectral(self, l, v, out, bufs, B):
        """out = K_l(v) (overwrites out); keeps X for the backward pass. With out None
        the last step (the inverse x transform) is left to the caller, from Z."""
        n, C, m = self.n, self.C, self.m
        st = self._stream()
        T1, P, X, Y, U, Z = bufs["T1"], bufs["P"], bufs["X"][l], bufs["Y"], bufs["U"], bufs["Z"]
        if self.fused_layers:
            _native.check(self.lib.swno_x_transform(B * n, 2 * m, n, C, _ptr(self.Fx), n, 1,
                                                    _ptr(v), _ptr(T1), _is_bf16(v), st))
        else:
            self.gemm(2 * m, C, n, B * n, self.Fx, (n, 1, 0), v, (C, 1, n * C), T1,
                      (C, 1, 2 * m * C))
        self.gemm(2 * m, 2 * m * C, n, B, self.Fy, (n, 1, 0), T1, (2 * m * C, 1, n * 2 * m * C),
                  P, (2 * m * C, 1, 4 * m * m * C))
        _native.check(self.lib.swno_spectral_pack(0, B, m, C, _ptr(P), _ptr(X), st))
        mm, w2 = m * m, 2 * C
        self.gemm(B, w2, w2, mm, X, (mm * w2, 1, w2), self.rm[l], (w2, 1, w2 * w2), Y,
                  (mm * w2, 1, w2))
        _native.check(self.lib.swno_spectral_pack(2, B, m, C, _ptr(Y), _ptr(U), st))
        self.gemm(n, 2 * m * C, 2 * m, B, self.Gy, (2 * m, 1, 0), U,
                  (2 * m * C, 1, 4 * m * m * C), Z, (2 * m * C, 1, n * 2 * m * C))
        if out is not None:
            self.gemm(n, C, 2 * m, B * n, self.Gx, (2 * m, 1, 0), Z, (C, 1, 2 * m * C), out,
                      (C, 1, n * C))

    def _spectral_backward(self, l, dpre, dv, bufs, B):
        """dv += K_l^T(dpre); writes the gradients of R_l. With dv None the last step
        (dv += Fx^T dT1) is left to the caller, from T1."""
        n, C, m = self.n, self.C, self.m
        st = self._stream()
        X, Y, U, Z, T1, P = bufs["X"][l], bufs["Y"], bufs["U"], bufs["Z"], bufs["T1"], bufs["P"]
        # dZ[b,y] = Gx^T dpre[b,y]
        if self.fused_layers:
            _native.check(self.lib.swno_x_transform(B * n, 2 * m, n, C, _ptr(self.Gx), 1, 2 * m,
                                                    _ptr(dpre), _ptr(Z), 0, st))
        else:
            self.gemm(2 * m, C, n, B * n, self.Gx, (1, 2 * m, 0), dpre, (C, 1, n * C), Z,
                      (C, 1, 2 * m * C))
        # dU[b] = Gy^T dZ[b]
        self.gemm(2 * m, 2 * m * C, n, B, self.Gy, (1, 2 * m, 0), Z,
                  (2 * m * C, 1, n * 2 * m * C), U, (2 * m * C, 1, 4 * m * m * C))
        _native.check(self.lib.swno_spectral_pack(3, B, m, C, _ptr(U), _ptr(Y), st))  # dY
        mm, w2 = m * m, 2 * C
        # dRm[mode] = X[:, mode]^T dY[:, mode]
        self.gemm(w2, w2, B, mm, X, (1, mm * w2, w2), Y, (mm * w2, 1, w2), self.drm,
                  (w2, 1, w2 * w2))
        _native.check(self.lib.swno_mix_assemble_backward(
            mm, C, _ptr(self.drm), _ptr(self.g[f"R{l}_re"]), _ptr(self.g[f"R{l}_im"]), 0, st))
        # dX[:, mode] = dY[:, mode] Rm^T  (into P's storage, which is free now)
        dX = P[: B * mm * w2]
        self.gemm(B, w2, w2, mm, Y, (mm * w2, 1, w2), self.rm[l], (1, w2, w2 * w2), dX,
                  (mm * w2, 1, w2))
        _native.check(self.lib.swno_spectral_pack(1, B, m, C, _ptr(dX), _ptr(U), st))  # dP
        # dT1[b] = Fy^T dP[b]
        self.gemm(n, 2 * m * C, 2 * m, B, self.Fy, (1, n, 0), U, (2 * m * C, 1, 4 * m * m * C),
                  T1, (2 * m * C, 1, n * 2 * m * C))
        # dv[b,y] += Fx^T dT1[b,y]
        if dv is not None:
            self.gemm(n, C, 2 * m, B * n, self.Fx, (1, n, 0), T1, (C, 1, 2 * m * C), dv,
                      (C, 1, n * C), beta=1.0)

    # --------------------------------------------------------------- public
    def forward(self, inp):
        """``inp`` (B, N, N, in_ch) float32 on the device -> (B, N, N) prediction."""
        B = inp.shape[0]
        if tuple(inp.shape[1:]) != (self.n, self.n, self.Cin):
            raise ValueError(f"input must be (B, {self.n}, {self.n}, {self.Cin})")
        bufs = self._buffers(B)
        st = self._stream()
        lib = self.lib
        n, C, H = self.n, self.C, self.H
        rows = B * n * n
        bufs["inp"].copy_(inp.reshape(rows, self.Cin))
        for l in range(self.L):
            _native.check(lib.swno_mix_assemble(self.m * self.m, C, _ptr(self.p[f"R{l}_re"]),
                                                _ptr(self.p[f"R{l}_im"]), _ptr(self.rm[l]), st))
        v = bufs["v"]
        # lifting: into v[0] itself, or (bf16 layer outputs) through the fp32 buffer of
        # layer 0's pre-activation, which layer 0 overwrites later
        lift = bufs["pre"][0] if _is_bf16(v[0]) else v[0]
        self.gemm(rows, C, self.Cin, 1, bufs["inp"], (self.Cin, 1, 0), self.p["P"], (C, 1, 0),
                  lift, (C, 1, 0))
        _native.check(lib.swno_bias_act(rows, C, _ptr(lift), _ptr(self.p["bp"]), _ptr(lift),
                                        _ptr(v[0]), 0, _is_bf16(v[0]), st))
        for l in range(self.L):
            pre = bufs["pre"][l]
            if self.fused_layers:
                self._spectral(l, v[l], None, bufs, B)
                _native.check(lib.swno_layer_forward(
                    B * n, n, self.m, C, _ptr(self.Gx), _ptr(bufs["Z"]), _ptr(v[l]),
                    _ptr(self.p[f"W{l}"]), _ptr(self.p[f"b{l}"]), _ptr(pre), _ptr(v[l + 1]),
                    int(l < self.L - 1), _is_bf16(v[l]) | (_is_bf16(v[l + 1]) << 1), st))
                continue
            self._spectral(l, v[l], pre, bufs, B)
            self.gemm(rows, C, C, 1, v[l], (C, 1, 0), self.p[f"W{l}"], (C, 1, 0), pre, (C, 1, 0),
                      beta=1.0)
            _native.check(lib.swno_bias_act(rows, C, _ptr(pre), _ptr(self.p[f"b{l}"]), _ptr(pre),
                                            _ptr(v[l + 1]), int(l < self.L - 1), 0, st))
        if self.fused_head:
            _native.check(lib.swno_head_forward(
                rows, C, H, _ptr(v[self.L]), _ptr(self.p["Q1"]), _ptr(self.p["bq1"]),
                _ptr(self.p["Q2"]), _ptr(self.p["bq2"]), _ptr(bufs["hpre"]), _ptr(bufs["out"]),
                _is_bf16(v[self.L]), st))
        else:
            self.gemm(rows, H, C, 1, v[self.L], (C, 1, 0), self.p["Q1"], (H, 1, 0),
                      bufs["hpre"], (H, 1, 0))
            _native.check(lib.swno_bias_act(rows, H, _ptr(bufs["hpre"]), _ptr(self.p["bq1"]),
                                            _ptr(bufs["hpre"]), _ptr(bufs["h"]), 1, 0, st))
            self.gemm(rows, 1, H, 1, bufs["h"], (H, 1, 0), self.p["Q2"], (1, 1, 0),
                      bufs["outpre"], (1, 1, 0))
            _native.check(lib.swno_bias_act(rows, 1, _ptr(bufs["outpre"]), _ptr(self.p["bq2"]),
                                            _ptr(bufs["out"]), _ptr(bufs["out"]), 0, 0, st))
        self._last_B = B
        return bufs["out"].view(B, n, n)

    def loss_and_backward(self, target, mask=None):
        """Masked MSE against ``target`` (B, N, N) of the last forward; fills ``grad_flat``.

        Returns the loss as a device scalar (no host synchronisation)."""
        B = self._last_B
        bufs = self._buffers(B)
        st = self._stream()
        lib = self.lib
        n, C, H = self.n, self.C, self.H
        rows = B * n * n
        t = target.reshape(rows).contiguous()
        mk = None if mask is None else mask.reshape(rows).contiguous()
        _native.check(lib.swno_mse(rows, _ptr(bufs["out"]), _ptr(t), _ptr(mk), _ptr(bufs["loss"]),
                                   _ptr(bufs["dout"]), _ptr(bufs["msews"]), st))
        self.backward_from(bufs["dout"], B)
        return bufs["loss"][0]

    def backward_from(self, dout, B):
        """Gradients of every parameter for d loss / d out = ``dout`` (rows, 1)."""
        bufs = self._buffers(B)
        st = self._stream()
        lib = self.lib
        n, C, H = self.n, self.C, self.H
        rows = B * n * n
        wgws = bufs["wgws"]
        g = self.g
        dv, dpre = bufs["dv"], bufs["dpre"]
        # projection
        if self.fused_head:
            _native.check(lib.swno_head_backward(
                rows, C, H, _ptr(dout), _ptr(bufs["hpre"]), _ptr(bufs["v"][self.L]),
                _ptr(self.p["Q1"]), _ptr(self.p["Q2"]), _ptr(dv), _ptr(g["Q1"]), _ptr(g["bq1"]),
                _ptr(g["Q2"]), _ptr(g["bq2"]), _ptr(bufs["headws"]), _is_bf16(bufs["v"][self.L]),
                st))
        else:
            self.weight_grad(rows, bufs["h"], H, dout, 1, g["Q2"], g["bq2"], wgws)
            self.gemm(rows, H, 1, 1, dout, (1, 1, 0), self.p["Q2"], (1, 1, 0), bufs["dh"],
                      (H, 1, 0))
            _native.check(lib.swno_gelu_backward(rows * H, _ptr(bufs["hpre"]),
                                                 _ptr(bufs["dh"]), _ptr(bufs["dh"]), st))
            self.weight_grad(rows, bufs["v"][self.L], C, bufs["dh"], H, g["Q1"], g["bq1"],
                             wgws)
            self.gemm(rows, C, H, 1, bufs["dh"], (H, 1, 0), self.p["Q1"], (1, H, 0), dv,
                      (C, 1, 0))
        # spectral layers, last to first
        if self.fused_layers:
            # dpre of the last layer is dv (no GELU after it); each fused kernel turns a
            # layer's dpre into the previous layer's dpre (or, for layer 0, into dv)
            cur, spare = dv, dpre
            for l in range(self.L - 1, -1, -1):
                self.weight_grad(rows, bufs["v"][l], C, cur, C, g[f"W{l}"], g[f"b{l}"], wgws)
                self._spectral_backward(l, cur, None, bufs, B)
                prev = bufs["pre"][l - 1] if l > 0 else None
                _native.check(lib.swno_layer_backward(
                    B * n, n, self.m, C, _ptr(cur), _ptr(self.p[f"W{l}"]), _ptr(self.Fx),
                    _ptr(bufs["T1"]), _ptr(prev), _ptr(spare), st))
                cur, spare = spare, cur
            self.weight_grad(rows, bufs["inp"], self.Cin, cur, C, g["P"], g["bp"], wgws)
            return
        for l in range(self.L - 1, -1, -1):
            v, pre = bufs["v"][l], bufs["pre"][l]
            if l < self.L - 1:
                _native.check(lib.swno_gelu_backward(rows * C, _ptr(pre), _ptr(dv), _ptr(dpre),
                                                     st))
            else:
                dpre.copy_(dv)
            self.weight_grad(rows, v, C, dpre, C, g[f"W{l}"], g[f"b{l}"], wgws)
            self.gemm(rows, C, C, 1, dpre, (C, 1, 0), self.p[f"W{l}"], (1, C, 0), dv, (C, 1, 0))
            self._spectral_backward(l, dpre, dv, bufs, B)
        # lifting
        self.weight_grad(rows, bufs["inp"], self.Cin, dv, C, g["P"], g["bp"], wgws)

    def adamw_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=1e-4):
        self.adam_step += 1
        _native.check(self.lib.swno_adamw(self.flat.numel(), _ptr(self.flat),
                                          _ptr(self.grad_flat), _ptr(self.adam_m),
                                          _ptr(self.adam_v), lr, beta1, beta2, eps, weight_decay,
                                          self.adam_step, self._stream()))

    def state_dict(self):
        return {k: self.p[k].detach().cpu().numpy().astype(np.float64) for k in self.names}

    def gradients(self):
        return {k: self.g[k].detach().cpu().numpy().astype(np.float64) for k in self.names}
