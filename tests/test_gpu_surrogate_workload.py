"""The learned surrogate at its workload sizes on the device: a 128 x 128 grid
at 1 mm, a 32-element ring of radius 50 mm, a 150 kHz Ricker, 1200 steps,
all 32 shots; the spectral operator with lifting to width 32, 4 layers,
16 x 16 retained modes and a projection through width 128.

* The operator's forward pass and every parameter gradient at 128^2 with 16
  x 16 modes (batch 4) against the float64 FFT-based oracle (bf16 operands:
  forward rel-L2 <= 2e-2, gradients <= 3e-2, as at the small size).
* The first FWI gradients the training pairs are made of (batched over two
  phantoms, all 32 shots, 1200 steps) against the float64 oracle's gradient
  from the same observed data: rel-L2 <= 5e-4 inside the ring.
* What the trained network gains: trained on 512 FDTD-generated pairs, its
  held-out speed RMSE inside the ring (64 phantoms) is below 0.85 of the
  homogeneous start's.
"""

import numpy as np
import pytest
import torch

import oracle
from oracle import surrogate_oracle as so
from soundwavesimulation_amd.surrogate.data import (BatchedFirstGradients, Geometry, generate,
                                                    speed_from_target)
from soundwavesimulation_amd.surrogate.operator import SpectralOperator, init_params
from soundwavesimulation_amd.surrogate.train import predict, rmse_inside, train
from soundwavesimulation_amd.torch_backend import simulate_shots_batch

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_operator_at_workload_size_matches_oracle():
    n, m, C, H, L, B = 128, 16, 32, 128, 4, 4
    print(f"operator n={n} modes={m} width={C} hidden={H} batch={B}", flush=True)
    rng = np.random.default_rng(17)
    params = init_params(seed=5, width=C, modes=m, layers=L, hidden=H)
    for l in range(L):
        params[f"R{l}_re"] = rng.standard_normal(params[f"R{l}_re"].shape) * 0.2
        params[f"R{l}_im"] = rng.standard_normal(params[f"R{l}_im"].shape) * 0.2
        params[f"b{l}"] = rng.standard_normal(C) * 0.1
    inp = rng.standard_normal((B, n, n, 3))
    target = rng.standard_normal((B, n, n))
    mask = (rng.random((B, n, n)) > 0.2).astype(float)
    op = SpectralOperator(params, n, DEV, width=C, modes=m, layers=L, hidden=H)
    assert op.fused_head and op.fused_layers
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


def test_first_gradients_at_workload_geometry_match_oracle():
    geo = Geometry()
    assert (geo.n, geo.n_elements, geo.n_steps) == (128, 32, 1200)
    cs = [geo.phantom(s) for s in (101, 102)]
    bg = BatchedFirstGradients(geo, DEV, 2)
    grads = bg.first_gradients(cs).double().cpu().numpy()
    gen = bg.base
    S = geo.n_elements
    r0, c0i, w = geo.array.footprint()
    m0 = 1.0 / geo.c0 ** 2
    for k, c in enumerate(cs):
        gen.solver_true.set_c(c)
        obs, _ = simulate_shots_batch(gen.solver_true, gen.sampler, gen.pulse, gen.rows_np,
                                      gen.cols_np, geo.n_steps)
        J64, g64, _ = oracle.gradient(np.full((geo.n, geo.n), geo.c0), geo.spacing, geo.spacing,
                                      geo.dt, gen.rows_np, gen.cols_np, geo.pulse, geo.n_steps,
                                      r0, c0i, w, obs.double().cpu().numpy(), np.arange(S))
        g_m = -g64 / (m0 * m0)
        e = _rel(grads[k][geo.mask], g_m[geo.mask])
        print(f"phantom {k}: first gradient rel-L2 vs float64 oracle {e:.2e}", flush=True)
        assert e <= 5e-4


def test_training_on_fdtd_pairs_beats_the_homogeneous_start():
    geo = Geometry()
    n_train, n_test = 512, 64
    print(f"generating {n_train + n_test} pairs at the workload geometry", flush=True)
    inputs, targets, t_gen, _ = generate(geo, DEV, n_train + n_test, seed0=7_000_000)
    print(f"  {t_gen:.1f} s", flush=True)
    op = SpectralOperator(init_params(seed=3), geo.n, DEV)
    history, t_train, steps = train(op, inputs, targets, geo.mask, n_train, 40, batch=64,
                                    lr=2e-3, seed=3, log_every=1e9)
    pred = predict(op, inputs[n_train:], batch=64)
    c_pred = speed_from_target(pred.cpu().numpy().astype(float))
    c_test = speed_from_target(targets[n_train:].cpu().numpy().astype(float))
    rm_net = np.mean([rmse_inside(c_pred[i], c_test[i], geo.mask) for i in range(n_test)])
    rm_h = np.mean([rmse_inside(np.full_like(c_test[i], geo.c0), c_test[i], geo.mask)
                    for i in range(n_test)])
    print(f"  trained {steps} steps in {t_train:.1f} s: loss {history[0]:.4f} -> "
          f"{history[-1]:.4f}; held-out RMSE network {rm_net:.2f} m/s, homogeneous start "
          f"{rm_h:.2f} m/s", flush=True)
    assert history[-1] < history[0]
    assert rm_net < 0.85 * rm_h
