"""GPU checks of the spectral neural operator kernels (libswno).

* GEMM: against the float64 product of the bf16-rounded operands (the only
  remaining error is fp32 accumulation, rel <= 1e-5), over transposed views,
  batches, split-K, beta accumulation and sizes that are not tile multiples.
* Elementwise kernels against float64 references.
* The operator's forward pass and every parameter gradient against the float64
  FFT-based oracle (bf16 operands: rel-L2 <= 2e-2 stated per quantity).
* A few AdamW steps reduce the training loss.
"""

import numpy as np
import pytest
import torch

from oracle import surrogate_oracle as so
from soundwavesimulation_amd.surrogate import _native
from soundwavesimulation_amd.surrogate.operator import SpectralOperator, init_params

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _bf16(x):
    return x.to(torch.bfloat16).to(torch.float64)


def _gemm(M, N, K, batch, A, a, B, b, C, c, alpha=1.0, beta=0.0, k_split=1):
    lib = _native.load_library()
    g = _native.GemmArgs(M, N, K, batch, k_split, A.data_ptr(), *a, B.data_ptr(), *b,
                         C.data_ptr(), *c, alpha, beta)
    ws = None
    need = int(lib.swno_gemm_workspace_floats(g))
    if need:
        ws = torch.empty(need, device=DEV)
    _native.check(lib.swno_gemm(g, None if ws is None else ws.data_ptr(),
                                torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("M,N,K,batch,ta,tb,k_split,beta", [
    (64, 64, 64, 1, False, False, 1, 0.0),
    (37, 50, 45, 3, False, False, 1, 0.0),     # ragged tiles and K
    (130, 96, 200, 2, True, False, 1, 1.0),    # transposed A, accumulate
    (33, 70, 31, 1, False, True, 1, 0.5),      # transposed B
    (32, 32, 20000, 1, True, False, 7, 0.0),   # split-K reduction
    (8, 1, 3, 5, False, False, 1, 0.0),
])
def test_gemm_matches_bf16_rounded_product(M, N, K, batch, ta, tb, k_split, beta):
    print(f"gemm M={M} N={N} K={K} batch={batch} ta={ta} tb={tb} split={k_split}", flush=True)
    g = torch.Generator().manual_seed(M * 1000 + N + K)
    A = torch.randn((batch, K, M) if ta else (batch, M, K), generator=g).to(DEV)
    B = torch.randn((batch, N, K) if tb else (batch, K, N), generator=g).to(DEV)
    C0 = torch.randn((batch, M, N), generator=g).to(DEV)
    C = C0.clone()
    a = (1, M, K * M) if ta else (K, 1, M * K)
    b = (1, K, N * K) if tb else (N, 1, K * N)
    _gemm(M, N, K, batch, A, a, B, b, C, (N, 1, M * N), alpha=0.75, beta=beta, k_split=k_split)
    torch.cuda.synchronize()
    Am = A.transpose(1, 2) if ta else A
    Bm = B.transpose(1, 2) if tb else B
    ref = 0.75 * torch.matmul(_bf16(Am.cpu()), _bf16(Bm.cpu())) + beta * C0.cpu().double()
    err = float((C.cpu().double() - ref).norm() / ref.norm())
    assert err <= 1e-5, err


@pytest.mark.parametrize("rows,M,N,bias,offset", [
    (70000, 32, 32, True, 0),      # a layer's W_l / b_l
    (70000, 32, 128, True, 0),     # Q1 / bq1
    (70000, 128, 1, True, 0),      # Q2 / bq2 (scalar B)
    (70000, 3, 32, True, 0),       # lifting P / bp (scalar A)
    (1000, 20, 24, False, 0),      # padded tiles, one partial per 64-row slab
    (1000, 20, 16, True, 0),
    (33, 16, 64, False, 0),        # a single partial slab, no column sums
    (5000, 8, 8, True, 1),         # misaligned bases: scalar loads on both sides
    (2_200_000, 32, 32, True, 0),  # more slabs than partials
])
def test_weight_grad_matches_bf16_rounded_product(rows, M, N, bias, offset):
    print(f"weight_grad rows={rows} M={M} N={N} bias={bias} offset={offset}", flush=True)
    lib = _native.load_library()
    g = torch.Generator().manual_seed(rows + M + N)
    A = torch.randn(rows * M + offset, generator=g).to(DEV)[offset:].view(rows, M)
    B = torch.randn(rows * N + offset, generator=g).to(DEV)[offset:].view(rows, N)
    dW = torch.full((M, N), float("nan"), device=DEV)
    db = torch.full((N,), float("nan"), device=DEV)
    ws = torch.empty(int(lib.swno_weight_grad_workspace_floats(rows, M, N)), device=DEV)
    st = torch.cuda.current_stream().cuda_stream
    args = (rows, M, N, A.data_ptr(), M, B.data_ptr(), N, dW.data_ptr(),
            db.data_ptr() if bias else None, ws.data_ptr(), 0, st)
    _native.check(lib.swno_weight_grad(*args))
    first = (dW.clone(), db.clone())
    _native.check(lib.swno_weight_grad(*args))  # fixed summation order: bitwise repeatable
    torch.cuda.synchronize()
    assert torch.equal(first[0], dW)
    ref = _bf16(A.cpu()).T @ _bf16(B.cpu())
    err = float((dW.cpu().double() - ref).norm() / ref.norm())
    assert err <= 1e-5, err
    if bias:
        assert torch.equal(first[1], db)
        ref_b = B.cpu().double().sum(0)
        errb = float((db.cpu().double() - ref_b).norm() / B.cpu().double().abs().sum(0).norm())
        assert errb <= 1e-6, errb
    else:
        assert torch.isnan(db).all()  # untouched


@pytest.mark.parametrize("slices,M,n,transposed", [(300, 32, 128, False), (300, 32, 128, True),
                                                   (77, 16, 32, False), (5, 20, 48, True),
                                                   (5000, 32, 32, False)])  # several per workgroup
def test_x_transform_matches_bf16_rounded_product(slices, M, n, transposed):
    print(f"x_transform slices={slices} M={M} n={n} transposed={transposed}", flush=True)
    lib = _native.load_library()
    g = torch.Generator().manual_seed(slices + M + n)
    A = torch.randn((n, M) if transposed else (M, n), generator=g).to(DEV)
    X = torch.randn((slices, n, 32), generator=g).to(DEV)
    out = torch.full((slices, M, 32), float("nan"), device=DEV)
    a_rs, a_cs = (1, M) if transposed else (n, 1)
    _native.check(lib.swno_x_transform(slices, M, n, 32, A.data_ptr(), a_rs, a_cs, X.data_ptr(),
                                       out.data_ptr(), 0, torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    Am = A.T if transposed else A
    ref = torch.matmul(_bf16(Am.cpu()), _bf16(X.cpu()))
    err = float((out.cpu().double() - ref).norm() / ref.norm())
    assert err <= 1e-5, err


@pytest.mark.parametrize("rows", [1000, 70001])  # a partial 64-row block; several per part
def test_head_kernels_match_float64(rows):
    print(f"head forward / backward rows={rows}", flush=True)
    lib = _native.load_library()
    st = torch.cuda.current_stream().cuda_stream
    g = torch.Generator().manual_seed(rows)
    C, H = 32, 128
    v = torch.randn((rows, C), generator=g).to(DEV)
    Q1 = (torch.randn((C, H), generator=g) * 0.3).to(DEV)
    bq1 = (torch.randn(H, generator=g) * 0.3).to(DEV)
    Q2 = (torch.randn(H, generator=g) * 0.2).to(DEV)
    bq2 = torch.randn(1, generator=g).to(DEV)
    dout = torch.randn(rows, generator=g).to(DEV)
    hpre = torch.empty((rows, H), device=DEV)
    out = torch.empty(rows, device=DEV)
    _native.check(lib.swno_head_forward(rows, C, H, v.data_ptr(), Q1.data_ptr(), bq1.data_ptr(),
                                        Q2.data_ptr(), bq2.data_ptr(), hpre.data_ptr(),
                                        out.data_ptr(), 0, st))
    dv = torch.empty((rows, C), device=DEV)
    dQ1, dbq1 = torch.empty((C, H), device=DEV), torch.empty(H, device=DEV)
    dQ2, dbq2 = torch.empty(H, device=DEV), torch.empty(1, device=DEV)
    ws = torch.empty(int(lib.swno_head_backward_workspace_floats(rows)), device=DEV)
    _native.check(lib.swno_head_backward(rows, C, H, dout.data_ptr(), hpre.data_ptr(),
                                         v.data_ptr(), Q1.data_ptr(), Q2.data_ptr(),
                                         dv.data_ptr(), dQ1.data_ptr(), dbq1.data_ptr(),
                                         dQ2.data_ptr(), dbq2.data_ptr(), ws.data_ptr(), 0, st))
    torch.cuda.synchronize()
    Q264 = Q2.cpu().double()
    hp64 = _bf16(v.cpu()) @ _bf16(Q1.cpu()) + bq1.cpu().double()
    assert _rel(hpre.cpu().double().numpy(), hp64.numpy()) <= 1e-5
    hp = hpre.cpu().double().numpy()  # the backward reads the kernel's own hpre
    out64 = so.gelu(hp) @ Q264.numpy() + float(bq2)
    assert _rel(out.cpu().double().numpy(), out64) <= 1e-5
    d = dout.cpu().double().numpy()
    dh = d[:, None] * Q264.numpy()[None, :] * so.gelu_grad(hp)
    dh_b = _bf16(torch.as_tensor(dh, dtype=torch.float32)).numpy()
    assert _rel(dv.cpu().double().numpy(), dh_b @ _bf16(Q1.cpu()).numpy().T) <= 1e-3
    assert _rel(dQ1.cpu().double().numpy(), _bf16(v.cpu()).numpy().T @ dh_b) <= 1e-3
    assert _rel(dbq1.cpu().double().numpy(), dh.sum(0)) <= 1e-5
    assert _rel(dQ2.cpu().double().numpy(), so.gelu(hp).T @ d) <= 1e-5
    assert abs(float(dbq2) - d.sum()) <= 1e-5 * np.abs(d).sum()
    assert not torch.isnan(dv).any()


def test_weight_grad_rejects_bad_shapes():
    lib = _native.load_library()
    x = torch.zeros(1024, device=DEV)
    st = torch.cuda.current_stream().cuda_stream
    p = x.data_ptr()
    assert lib.swno_weight_grad(16, 129, 4, p, 129, p, 4, p, None, p, 0, st) != 0  # M > 128
    assert lib.swno_weight_grad(16, 33, 4, p + 4, 36, p, 4, p, None, p, 0, st) != 0  # wide scalar
    assert lib.swno_weight_grad(16, 4, 12, p, 4, p, 12, p, p, p, 0, st) != 0  # 12 / 4 ∤ 256
    assert lib.swno_weight_grad(16, 4, 4, p, 2, p, 4, p, None, p, 0, st) != 0  # lda < M


def test_elementwise_kernels():
    lib = _native.load_library()
    st = torch.cuda.current_stream().cuda_stream
    g = torch.Generator().manual_seed(5)
    rows, cols = 5000, 24
    x = torch.randn((rows, cols), generator=g).to(DEV)
    bias = torch.randn(cols, generator=g).to(DEV)
    pre = torch.empty_like(x)
    y = torch.empty_like(x)
    _native.check(lib.swno_bias_act(rows, cols, x.data_ptr(), bias.data_ptr(), pre.data_ptr(),
                                    y.data_ptr(), 1, 0, st))
    dy = torch.randn((rows, cols), generator=g).to(DEV)
    dx = torch.empty_like(dy)
    _native.check(lib.swno_gelu_backward(rows * cols, pre.data_ptr(), dy.data_ptr(),
                                         dx.data_ptr(), st))
    ws = torch.empty(int(lib.swno_col_sum_workspace_floats(rows, cols)), device=DEV)
    out = torch.ones(cols, device=DEV)
    _native.check(lib.swno_col_sum(rows, cols, dy.data_ptr(), 2.0, out.data_ptr(),
                                   ws.data_ptr(), st))
    torch.cuda.synchronize()
    pre64 = x.cpu().double() + bias.cpu().double()
    np.testing.assert_allclose(pre.cpu().numpy(), pre64.numpy(), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(y.cpu().numpy(), so.gelu(pre64.numpy()), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(dx.cpu().numpy(), dy.cpu().double().numpy() *
                               so.gelu_grad(pre64.numpy()), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(out.cpu().numpy(), 2.0 + dy.cpu().double().sum(0).numpy(),
                               rtol=1e-5, atol=1e-4)
    # masked MSE
    n = 70000
    p = torch.randn(n, generator=g).to(DEV)
    t = torch.randn(n, generator=g).to(DEV)
    mk = (torch.rand(n, generator=g) > 0.4).float().to(DEV)
    loss = torch.empty(2, device=DEV)
    grad = torch.empty(n, device=DEV)
    ws = torch.empty(int(lib.swno_mse_workspace_floats(n)), device=DEV)
    _native.check(lib.swno_mse(n, p.data_ptr(), t.data_ptr(), mk.data_ptr(), loss.data_ptr(),
                               grad.data_ptr(), ws.data_ptr(), st))
    torch.cuda.synchronize()
    l64, g64 = so.mse(p.cpu().double().numpy(), t.cpu().double().numpy(),
                      mk.cpu().double().numpy())
    assert abs(float(loss[0]) - l64) <= 1e-5 * l64
    np.testing.assert_allclose(grad.cpu().numpy(), g64, rtol=1e-5, atol=1e-9)
    # AdamW: two steps against the closed form
    prm = torch.randn(1000, generator=g).to(DEV)
    p0 = prm.clone()
    gr = torch.randn(1000, generator=g).to(DEV)
    m = torch.zeros_like(prm)
    v = torch.zeros_like(prm)
    for step in (1, 2):
        _native.check(lib.swno_adamw(1000, prm.data_ptr(), gr.data_ptr(), m.data_ptr(),
                                     v.data_ptr(), 1e-2, 0.9, 0.999, 1e-8, 1e-2, step, st))
    torch.cuda.synchronize()
    pr = p0.cpu().double().numpy()
    gg = gr.cpu().double().numpy()
    mm = np.zeros_like(pr)
    vv = np.zeros_like(pr)
    for step in (1, 2):
        mm = 0.9 * mm + 0.1 * gg
        vv = 0.999 * vv + 0.001 * gg * gg
        pr = pr - 1e-2 * ((mm / (1 - 0.9 ** step)) / (np.sqrt(vv / (1 - 0.999 ** step)) + 1e-8)
                          + 1e-2 * pr)
    np.testing.assert_allclose(prm.cpu().numpy(), pr, rtol=1e-5, atol=1e-6)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("C,H", [(8, 16), (32, 128)])  # (32, 128): fused projection head
def test_operator_forward_and_gradients_match_oracle(C, H):
    n, m, L, B = 32, 8, 4, 3
    print(f"operator n={n} modes={m} width={C} hidden={H}", flush=True)
    rng = np.random.default_rng(7)
    params = init_params(seed=11, width=C, modes=m, layers=L, hidden=H)
    for l in range(L):
        params[f"R{l}_re"] = rng.standard_normal(params[f"R{l}_re"].shape) * 0.2
        params[f"R{l}_im"] = rng.standard_normal(params[f"R{l}_im"].shape) * 0.2
        params[f"b{l}"] = rng.standard_normal(C) * 0.1
    inp = rng.standard_normal((B, n, n, 3))
    target = rng.standard_normal((B, n, n))
    mask = (rng.random((B, n, n)) > 0.2).astype(float)
    op = SpectralOperator(params, n, DEV, width=C, modes=m, layers=L, hidden=H)
    out = op.forward(torch.as_tensor(inp, dtype=torch.float32, device=DEV))
    loss = op.loss_and_backward(torch.as_tensor(target, dtype=torch.float32, device=DEV),
                                torch.as_tensor(mask, dtype=torch.float32, device=DEV))
    torch.cuda.synchronize()
    ref, saved = so.forward(params, inp, modes=m, layers=L)
    l64, dout = so.mse(ref, target, mask)
    ref_g = so.backward(params, saved, dout, modes=m, layers=L)
    e_out = _rel(out.cpu().numpy(), ref)
    print(f"  forward rel-L2 {e_out:.2e}; loss {float(loss):.6f} vs {l64:.6f}")
    assert e_out <= 2e-2
    assert abs(float(loss) - l64) <= 2e-2 * l64
    got = op.gradients()
    for k in params:
        e = _rel(got[k], ref_g[k])
        print(f"  grad {k}: rel-L2 {e:.2e}")
        assert e <= 3e-2, k


@pytest.mark.parametrize("switch,B", [("fused_head", 5), ("fused_layers", 5),
                                      ("fused_layers", 300)])  # 9600 slices: several per workgroup
def test_fused_kernels_match_separate_kernels(switch, B):
    n, m, C, H = 32, 8, 32, 128
    print(f"{switch}: fused vs separate GEMM / elementwise kernels, batch {B}", flush=True)
    rng = np.random.default_rng(3)
    params = init_params(seed=4, width=C, modes=m, hidden=H)
    params["bq1"] = rng.standard_normal(H) * 0.3
    params["bq2"] = rng.standard_normal(1)
    inp = torch.as_tensor(rng.standard_normal((B, n, n, 3)), dtype=torch.float32, device=DEV)
    target = torch.as_tensor(rng.standard_normal((B, n, n)), dtype=torch.float32, device=DEV)
    res = []
    for fused in (True, False):
        op = SpectralOperator(params, n, DEV, width=C, modes=m, hidden=H)
        assert op.fused_head and op.fused_layers
        setattr(op, switch, fused)
        out = op.forward(inp).clone()
        op.loss_and_backward(target)
        torch.cuda.synchronize()
        res.append((out.cpu().numpy(), op.gradients(), op))
    e = _rel(res[0][0], res[1][0])
    print(f"  out rel-L2 {e:.2e}")
    assert e <= 5e-3
    for k in params:
        e = _rel(res[0][1][k], res[1][1][k])
        print(f"  grad {k}: rel-L2 {e:.2e}")
        assert e <= 1e-2, k
    # repeatable bit for bit (fixed-order sums)
    op = res[0][2]
    op.forward(inp)
    op.loss_and_backward(target)
    torch.cuda.synchronize()
    again = op.gradients()
    for k in params:
        assert np.array_equal(again[k], res[0][1][k]), k


def test_adamw_training_reduces_loss():
    n, m, C, H, L, B = 32, 8, 8, 16, 4, 8
    rng = np.random.default_rng(9)
    params = init_params(seed=2, width=C, modes=m, layers=L, hidden=H)
    op = SpectralOperator(params, n, DEV, width=C, modes=m, layers=L, hidden=H)
    y, x = np.meshgrid(np.linspace(-1, 1, n), np.linspace(-1, 1, n), indexing="ij")
    a = rng.standard_normal((B, 1, 1))
    field = a * np.sin(np.pi * x)[None] + 0.5 * a * np.cos(np.pi * y)[None]
    inp = np.stack([field, np.broadcast_to(x, field.shape), np.broadcast_to(y, field.shape)],
                   axis=-1)
    target = 2.0 * field
    ti = torch.as_tensor(inp, dtype=torch.float32, device=DEV)
    tt = torch.as_tensor(target, dtype=torch.float32, device=DEV)
    losses = []
    for _ in range(60):
        op.forward(ti)
        losses.append(float(op.loss_and_backward(tt)))
        op.adamw_step(3e-3, weight_decay=0.0)
    print(f"  loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.2 * losses[0]


def test_batched_first_gradients_equal_one_at_a_time():
    from soundwavesimulation_amd.surrogate.data import (BatchedFirstGradients,
                                                        FirstGradientGenerator, Geometry)

    geo = Geometry(n=64, n_elements=8, radius_m=25e-3, n_steps=300, f_peak_hz=200e3)
    print("batched first gradients, 3 phantoms", flush=True)
    cs = [geo.phantom(s) for s in range(3)]
    bg = BatchedFirstGradients(geo, DEV, 3)
    gb = bg.first_gradients(cs).clone()
    gen = FirstGradientGenerator(geo, DEV)
    for k, c in enumerate(cs):
        g1, _ = gen.first_gradient(c)
        torch.cuda.synchronize()
        assert float(g1.abs().max()) > 0.0
        err = float((gb[k] - g1).norm() / g1.norm())
        print(f"  phantom {k}: rel {err:.1e}")
        assert err <= 1e-6


@pytest.mark.parametrize("B", [5, 300])  # 300: 9600 slices, several per workgroup
def test_bf16_layer_outputs_give_the_same_bits(B, monkeypatch):
    """The fused path stores the layer outputs v_1 .. v_L as bf16: every kernel
    that reads them rounds them to bf16 for an MFMA product, so the output and
    every gradient must equal those of fp32 storage (SWNO_BF16_ACT=0) bit for
    bit."""
    n, m, C, H = 32, 8, 32, 128
    print(f"bf16 vs fp32 layer outputs, batch {B}", flush=True)
    rng = np.random.default_rng(5)
    params = init_params(seed=6, width=C, modes=m, hidden=H)
    for l in range(4):
        params[f"b{l}"] = rng.standard_normal(C) * 0.1
    inp = torch.as_tensor(rng.standard_normal((B, n, n, 3)), dtype=torch.float32, device=DEV)
    target = torch.as_tensor(rng.standard_normal((B, n, n)), dtype=torch.float32, device=DEV)
    res = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("SWNO_BF16_ACT", flag)
        op = SpectralOperator(params, n, DEV, width=C, modes=m, hidden=H)
        out = op.forward(inp).clone()
        loss = float(op.loss_and_backward(target))
        torch.cuda.synchronize()
        dt = op._buffers(B)["v"][1].dtype
        assert dt == (torch.bfloat16 if flag == "1" else torch.float32)
        res[flag] = (out, loss, op.gradients())
        del op
    assert torch.equal(res["0"][0], res["1"][0]) and res["0"][1] == res["1"][1]
    for k in params:
        assert np.array_equal(res["0"][2][k], res["1"][2][k]), k


def test_kernels_read_bf16_activations_like_their_rounded_fp32():
    """x_transform, the head kernels and weight_grad given bf16 activations
    equal the same kernels given the fp32 values of those codes, bit for bit
    (partial 64-row blocks and several slices per workgroup)."""
    print("bf16 activation inputs of the fused kernels", flush=True)
    lib = _native.load_library()
    st = torch.cuda.current_stream().cuda_stream
    g = torch.Generator().manual_seed(17)
    # x transform
    slices, M, n = 5000, 32, 32
    A = torch.randn((M, n), generator=g).to(DEV)
    Xb = torch.randn((slices, n, 32), generator=g).to(DEV).to(torch.bfloat16)
    Xf = Xb.float()
    outs = []
    for X, flag in ((Xf, 0), (Xb, 1)):
        o = torch.full((slices, M, 32), float("nan"), device=DEV)
        _native.check(lib.swno_x_transform(slices, M, n, 32, A.data_ptr(), n, 1, X.data_ptr(),
                                           o.data_ptr(), flag, st))
        outs.append(o)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    # head forward / backward and the weight gradient on v
    rows, C, H = 70001, 32, 128
    vb = torch.randn((rows, C), generator=g).to(DEV).to(torch.bfloat16)
    Q1 = (torch.randn((C, H), generator=g) * 0.3).to(DEV)
    bq1 = (torch.randn(H, generator=g) * 0.3).to(DEV)
    Q2 = (torch.randn(H, generator=g) * 0.2).to(DEV)
    bq2 = torch.randn(1, generator=g).to(DEV)
    dout = torch.randn(rows, generator=g).to(DEV)
    dpre = torch.randn((rows, C), generator=g).to(DEV)
    res = []
    for v, flag in ((vb.float(), 0), (vb, 1)):
        hpre = torch.empty((rows, H), device=DEV)
        out = torch.empty(rows, device=DEV)
        _native.check(lib.swno_head_forward(rows, C, H, v.data_ptr(), Q1.data_ptr(),
                                            bq1.data_ptr(), Q2.data_ptr(), bq2.data_ptr(),
                                            hpre.data_ptr(), out.data_ptr(), flag, st))
        dv = torch.empty((rows, C), device=DEV)
        grads = [torch.empty((C, H), device=DEV), torch.empty(H, device=DEV),
                 torch.empty(H, device=DEV), torch.empty(1, device=DEV)]
        ws = torch.empty(int(lib.swno_head_backward_workspace_floats(rows)), device=DEV)
        _native.check(lib.swno_head_backward(rows, C, H, dout.data_ptr(), hpre.data_ptr(),
                                             v.data_ptr(), Q1.data_ptr(), Q2.data_ptr(),
                                             dv.data_ptr(), *[t.data_ptr() for t in grads],
                                             ws.data_ptr(), flag, st))
        dW, db = torch.empty((C, C), device=DEV), torch.empty(C, device=DEV)
        wws = torch.empty(int(lib.swno_weight_grad_workspace_floats(rows, C, C)), device=DEV)
        _native.check(lib.swno_weight_grad(rows, C, C, v.data_ptr(), C, dpre.data_ptr(), C,
                                           dW.data_ptr(), db.data_ptr(), wws.data_ptr(), flag,
                                           st))
        res.append([hpre, out, dv, *grads, dW, db])
    torch.cuda.synchronize()
    for a, b in zip(*res):
        assert torch.equal(a, b)
