"""Train the spectral neural operator on FDTD-generated pairs and evaluate it.

``python -m soundwavesimulation_amd.surrogate.train --samples 4096 --epochs 60``
generates the pairs on the GPU (``data.generate``), trains with AdamW (batch
64, 3584 train / 512 test for 4096 samples), reports the test speed RMSE of
the network against the homogeneous start, and runs FWI on held-out phantoms
from the homogeneous start and from the network's prediction (falling back to
the homogeneous start when the prediction's data misfit is higher), both with
the same optimiser and iteration budget. One JSON line is printed (and
written to ``--out``).

FLOP accounting for the MFMA utilisation: every GEMM of a training step is
counted (2 M N K per product, forward and backward), divided by the step time
and by the MI355X dense bf16 peak (2.5 PFLOP/s).
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from ..inversion.fwi import TorchLeastSquaresFWI
from ..inversion.medium import c_from_m, m_from_c
from ..inversion.optim import polak_ribiere
from ..torch_backend import TorchFDTD2D, TorchScalarWave2D, simulate_shots_batch
from .data import Geometry, generate, speed_from_target
from .operator import SpectralOperator, init_params

BF16_PEAK_TFLOPS = 2500.0


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def step_flops(B, n, C=32, m=16, L=4, H=128, cin=3):
    """GEMM FLOPs of one forward + backward pass (2 M N K per product)."""
    P = B * n * n
    spec_fwd = (2 * (2 * m) * C * n * B * n + 2 * (2 * m) * (2 * m * C) * n * B
                + 2 * B * (2 * C) * (2 * C) * m * m + 2 * n * (2 * m * C) * (2 * m) * B
                + 2 * n * C * (2 * m) * B * n)
    fwd = 2 * P * cin * C + L * (spec_fwd + 2 * P * C * C) + 2 * P * C * H + 2 * P * H
    # backward: the data gradient of every product plus every weight gradient
    spec_bwd = spec_fwd + 2 * (2 * C) * (2 * C) * B * m * m
    bwd = (2 * P * cin * C + L * (spec_bwd + 4 * P * C * C) + 4 * P * C * H + 4 * P * H)
    return fwd + bwd


def rmse_inside(c_est, c_true, mask):
    d = (c_est - c_true)[mask]
    return float(np.sqrt(np.mean(d * d)))


def train(op, inputs, targets, mask, n_train, epochs, batch=64, lr=1e-3, weight_decay=1e-4,
          seed=0, log_every=20.0):
    dev = inputs.device
    n = op.n
    mask_b = torch.as_tensor(mask, dtype=torch.float32, device=dev).expand(batch, n, n)
    rng = np.random.default_rng(seed)
    history = []
    steps_per_epoch = n_train // batch
    total_steps = epochs * steps_per_epoch
    t0 = time.perf_counter()
    last = t0
    step = 0
    for ep in range(epochs):
        perm = torch.as_tensor(rng.permutation(n_train), device=dev)
        acc = torch.zeros((), device=dev)
        for k in range(steps_per_epoch):
            idx = perm[k * batch:(k + 1) * batch]
            xb = inputs.index_select(0, idx)
            yb = targets.index_select(0, idx)
            op.forward(xb)
            acc = acc + op.loss_and_backward(yb, mask_b)
            # cosine decay to 5 % of lr
            frac = step / max(total_steps - 1, 1)
            op.adamw_step(lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * frac))),
                          weight_decay=weight_decay)
            step += 1
        history.append(float(acc) / steps_per_epoch)
        now = time.perf_counter()
        if now - last > log_every or ep == epochs - 1:
            log(f"epoch {ep + 1}/{epochs}: train loss {history[-1]:.5f} "
                f"({now - t0:.1f} s)")
            last = now
    torch.cuda.synchronize()
    return history, time.perf_counter() - t0, step


def predict(op, inputs, batch=64):
    outs = []
    for k in range(0, inputs.shape[0], batch):
        xb = inputs[k:k + batch]
        if xb.shape[0] < batch:  # keep one buffer set: pad the last batch
            pad = torch.zeros((batch - xb.shape[0], *xb.shape[1:]), device=xb.device)
            out = op.forward(torch.cat([xb, pad]))[: xb.shape[0]].clone()
        else:
            out = op.forward(xb).clone()
        outs.append(out)
    return torch.cat(outs)


def fwi_runs(geometry, gen, c_true, c_pred, device, iterations, target_rmse):
    """Plain (homogeneous start) vs warm-started FWI on one phantom."""
    g = geometry
    S = g.n_elements
    gen.solver_true.set_c(c_true)
    traces, _ = simulate_shots_batch(gen.solver_true, gen.sampler, gen.pulse, gen.rows_np,
                                     gen.cols_np, g.n_steps)
    solver = TorchFDTD2D(TorchScalarWave2D(np.full((g.n, g.n), g.c0), g.spacing, device=device),
                         g.dt)
    problem = TorchLeastSquaresFWI(solver, gen.sampler, g.array, gen.pulse, g.n_steps, S, traces,
                                   g.mask, c_min=1400.0, c_max=1650.0, c_background=g.c0,
                                   parallel_shots=S)
    out = {}
    m_plain = m_from_c(np.full((g.n, g.n), g.c0))
    m_warm = m_from_c(np.clip(c_pred, 1400.0, 1650.0))
    m_warm[~g.mask] = m_from_c(g.c0)
    J_plain0 = problem.misfit(m_plain)
    J_warm0 = problem.misfit(m_warm)
    fallback = J_warm0 > J_plain0
    starts = {"plain": m_plain, "warm": m_plain if fallback else m_warm}
    for name, m0 in starts.items():
        rm, tm = [], []
        t0 = time.perf_counter()

        def mg(m):
            val, grad = problem.misfit_and_grad(m)
            rm.append(rmse_inside(c_from_m(problem.project(m)), c_true, g.mask))
            tm.append(time.perf_counter() - t0)
            return val, grad

        rm.append(rmse_inside(c_from_m(problem.project(m0)), c_true, g.mask))
        tm.append(0.0)
        m, hist = polak_ribiere(m0, problem.misfit, mg, problem.project, g.mask,
                                max_iter=iterations, verbose=False)
        t_hit = next((t for r, t in zip(rm, tm) if r <= target_rmse), None)
        out[name] = {"rmse_start": rm[0], "rmse_final": rmse_inside(c_from_m(m), c_true, g.mask),
                     "misfit_final": float(hist["misfit"][-1]), "seconds": tm[-1],
                     "seconds_to_target": t_hit, "iterations": len(hist["misfit"]) - 1}
    out["fallback_to_plain"] = bool(fallback)
    out["misfit_start"] = {"plain": float(J_plain0), "warm": float(J_warm0)}
    return out


def time_steps(n, batch, steps, device, seed=0, warmup=5):
    """Training steps (forward, masked MSE backward, AdamW) on random data, timed."""
    op = SpectralOperator(init_params(seed=seed), n, device)
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.randn((batch, n, n, 3), generator=g).to(device)
    y = torch.randn((batch, n, n), generator=g).to(device)
    for _ in range(warmup):
        op.forward(x)
        op.loss_and_backward(y)
        op.adamw_step(1e-3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        op.forward(x)
        op.loss_and_backward(y)
        op.adamw_step(1e-3)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    flops = step_flops(batch, n)
    return {"workload": f"operator_training_step_{n}x{n}_batch{batch}", "steps": steps,
            "ms": round(t * 1e3, 3), "gflop": round(flops / 1e9, 2),
            "tflops": round(flops / t / 1e12, 2),
            "mfma_frac_of_bf16_peak": round(flops / t / 1e12 / BF16_PEAK_TFLOPS, 5),
            "fused": bool(op.fused_head and op.fused_layers)}


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--samples", type=int, default=4096)
    p.add_argument("--test", type=int, default=512)
    p.add_argument("--epochs", type=int, default=60)
    p.add_argument("--batch", type=int, default=64)
    p.add_argument("--lr", type=float, default=2e-3)
    p.add_argument("--grid", type=int, default=128)
    p.add_argument("--steps", type=int, default=1200)
    p.add_argument("--fwi-phantoms", type=int, default=2)
    p.add_argument("--fwi-iterations", type=int, default=10)
    p.add_argument("--target-rmse", type=float, default=10.0)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--out", default=None)
    p.add_argument("--step-only", type=int, default=0, metavar="K",
                   help="only time K training steps on random data (no FDTD pairs)")
    a = p.parse_args(argv)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if a.step_only:
        print(json.dumps(time_steps(a.grid, a.batch, a.step_only, dev, seed=a.seed)), flush=True)
        return
    geo = Geometry(n=a.grid, n_steps=a.steps)
    log(f"generating {a.samples} pairs: {a.grid}^2, {geo.n_elements} shots, {a.steps} steps")
    inputs, targets, t_gen, gen = generate(geo, dev, a.samples, seed0=a.seed * 1_000_000, log=log)
    cells = 2.0 * a.grid * a.grid * a.steps * geo.n_elements * a.samples
    log(f"data: {t_gen:.1f} s ({a.samples / t_gen:.1f} pairs/s, "
        f"{cells / t_gen / 1e9:.1f} Gcell/s forward+adjoint)")
    n_train = a.samples - a.test
    op = SpectralOperator(init_params(seed=a.seed), a.grid, dev)
    # time a few steps for the MFMA figure
    xb, yb = inputs[:a.batch].contiguous(), targets[:a.batch].contiguous()
    mask_b = torch.as_tensor(geo.mask, dtype=torch.float32, device=dev).expand(a.batch, a.grid,
                                                                               a.grid)
    for _ in range(2):
        op.forward(xb)
        op.loss_and_backward(yb, mask_b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n_timed = 10
    for _ in range(n_timed):
        op.forward(xb)
        op.loss_and_backward(yb, mask_b)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / n_timed
    flops = step_flops(a.batch, a.grid)
    op = SpectralOperator(init_params(seed=a.seed), a.grid, dev)  # fresh weights
    history, t_train, n_steps = train(op, inputs, targets, geo.mask, n_train, a.epochs,
                                      batch=a.batch, lr=a.lr, seed=a.seed)
    pred = predict(op, inputs[n_train:], batch=a.batch)
    c_pred = speed_from_target(pred.cpu().numpy().astype(float))
    c_test = speed_from_target(targets[n_train:].cpu().numpy().astype(float))
    rm_net = [rmse_inside(c_pred[i], c_test[i], geo.mask) for i in range(a.test)]
    rm_h = [rmse_inside(np.full_like(c_test[i], geo.c0), c_test[i], geo.mask)
            for i in range(a.test)]
    log(f"test RMSE: network {np.mean(rm_net):.2f} m/s, homogeneous start {np.mean(rm_h):.2f}")
    fwi = []
    for i in range(min(a.fwi_phantoms, a.test)):
        r = fwi_runs(geo, gen, c_test[i], c_pred[i], dev, a.fwi_iterations, a.target_rmse)
        log(f"FWI phantom {i}: plain {r['plain']['rmse_final']:.2f} m/s in "
            f"{r['plain']['seconds']:.1f} s, warm {r['warm']['rmse_final']:.2f} m/s in "
            f"{r['warm']['seconds']:.1f} s (fallback {r['fallback_to_plain']})")
        fwi.append(r)
    report = {
        "workload": f"surrogate_{a.grid}x{a.grid}_{a.samples}pairs",
        "data": {"pairs": a.samples, "seconds": round(t_gen, 2),
                 "pairs_per_s": round(a.samples / t_gen, 2),
                 "fdtd_gcell_s": round(cells / t_gen / 1e9, 2)},
        "training": {"train": n_train, "test": a.test, "epochs": a.epochs, "batch": a.batch,
                     "steps": n_steps, "seconds": round(t_train, 2),
                     "loss_first_epoch": history[0], "loss_last_epoch": history[-1]},
        "operator_step": {"ms": round(t_step * 1e3, 3), "gflop": round(flops / 1e9, 2),
                          "tflops": round(flops / t_step / 1e12, 2),
                          "mfma_frac_of_bf16_peak": round(flops / t_step / 1e12 /
                                                          BF16_PEAK_TFLOPS, 5)},
        "test_rmse_m_s": {"network": round(float(np.mean(rm_net)), 3),
                          "homogeneous_start": round(float(np.mean(rm_h)), 3),
                          "network_median": round(float(np.median(rm_net)), 3)},
        "fwi": fwi,
    }
    line = json.dumps(report)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
