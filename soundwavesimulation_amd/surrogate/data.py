"""Training pairs for the surrogate, generated by the FDTD on the device.

One pair = (first FWI gradient image, true speed map) for a seeded ellipse
phantom: observed traces are simulated on the true medium (all shots of the
ring), the residual is taken against the traces of the homogeneous starting
medium, and the adjoint-state gradient with respect to slowness squared at
that start is the input image. The starting medium is the same for every
pair, so its forward run (with checkpoints) is done once; each pair costs
one forward simulation and one checkpointed adjoint pass.

Default geometry: 128 x 128 grid at 1 mm, 32-element ring of radius 50 mm,
150 kHz Ricker, 1200 steps, homogeneous 1500 m/s start, 2-10 ellipses with
speeds 1400-1650 m/s.
"""

from __future__ import annotations

import time

import numpy as np
import torch

from .. import _native
from ..array import RingArray
from ..experiment import axis_centers, stable_dt
from ..inversion.fwi import CheckpointedGradient, _i32, residual_and_misfit
from ..inversion.medium import interior_mask
from ..phantoms import ellipse_phantom
from ..sources import ricker
from ..torch_backend import (TorchFDTD2D, TorchRingSampler, TorchScalarWave2D,
                             _source_values_tensor, simulate_shots_batch)

_F32 = torch.float32


class Geometry:
    """Acquisition and grid of the surrogate's training problems."""

    def __init__(self, n=128, spacing_m=1e-3, n_elements=32, radius_m=50e-3, f_peak_hz=150e3,
                 n_steps=1200, c0=1500.0, c_max=1650.0, cfl=0.3, margin_cells=2):
        self.n, self.spacing, self.n_elements = int(n), float(spacing_m), int(n_elements)
        self.radius, self.f_peak, self.n_steps = float(radius_m), float(f_peak_hz), int(n_steps)
        self.c0, self.c_max = float(c0), float(c_max)
        self.x = axis_centers(self.n, self.spacing)
        self.dt = stable_dt(c_max, self.spacing, self.spacing, cfl=cfl)
        self.pulse = ricker(self.dt, self.f_peak)
        self.array = RingArray(self.n_elements, self.radius, self.x, self.x)
        self.mask = interior_mask(self.x, self.x, self.radius, margin_cells * self.spacing)

    def phantom(self, seed, n_min=2, n_max=10):
        rng = np.random.default_rng(seed)
        k = int(rng.integers(n_min, n_max + 1))
        return ellipse_phantom(self.n, self.n, self.spacing, k, seed + 1_000_003, self.radius)

    def coordinates(self):
        t = np.linspace(-1.0, 1.0, self.n)
        yy, xx = np.meshgrid(t, t, indexing="ij")
        return xx, yy


class FirstGradientGenerator:
    """Device generator of (first gradient, speed) pairs for one geometry."""

    def __init__(self, geometry, device):
        g = geometry
        self.g = g
        self.device = device
        S = g.n_elements
        c0 = np.full((g.n, g.n), g.c0)
        self.solver_true = TorchFDTD2D(TorchScalarWave2D(c0, g.spacing, device=device), g.dt)
        self.solver0 = TorchFDTD2D(TorchScalarWave2D(c0, g.spacing, device=device), g.dt)
        self.sampler = TorchRingSampler(g.array, device=device)
        rc = [g.array.inject_rows_cols(i) for i in range(S)]
        self.rows_np = np.array([r for r, _ in rc])
        self.cols_np = np.array([c for _, c in rc])
        self.pulse = torch.as_tensor(g.pulse, dtype=_F32, device=device)
        self.engine = CheckpointedGradient(self.solver0, self.sampler, S, g.n_steps)
        values = _source_values_tensor(self.pulse, g.n_steps, device)
        self.samples0 = self.engine.forward(_i32(self.rows_np, device), _i32(self.cols_np, device),
                                            values).clone()
        self.index = _i32(np.arange(S), device)
        self.m0 = 1.0 / g.c0 ** 2
        self.grad = torch.zeros((g.n, g.n), device=device, dtype=_F32)

    def first_gradient(self, c_true):
        """Gradient of the misfit w.r.t. slowness squared at the start (device, (n, n))."""
        self.solver_true.set_c(c_true)
        traces, _ = simulate_shots_batch(self.solver_true, self.sampler, self.pulse,
                                         self.rows_np, self.cols_np, self.g.n_steps)
        residual, misfit = residual_and_misfit(self.samples0, traces, self.index, self.index)
        self.grad.zero_()
        self.engine.backward(residual, self.grad, self.solver0.dt2)
        return -self.grad / (self.m0 * self.m0), misfit

    def make_input(self, grad_m):
        """Network input (n, n, 3): masked gradient scaled to max |.| = 1, x, y."""
        g = self.g
        mask = torch.as_tensor(g.mask, device=self.device)
        gm = torch.where(mask, grad_m, torch.zeros_like(grad_m))
        gm = gm / gm.abs().max().clamp_min(1e-30)
        xx, yy = g.coordinates()
        coords = torch.as_tensor(np.stack([xx, yy], axis=-1), dtype=_F32, device=self.device)
        return torch.cat([gm[..., None], coords], dim=-1)


class BatchedFirstGradients:
    """``batch`` phantoms at a time through batched launches.

    All phantoms share the homogeneous start, so each segment of the reverse
    pass recomputes the start's forward field once and the adjoints of all
    phantoms (their own residuals, states and gradients) run as one batched
    launch per tick against it; the observed data of all phantoms are one
    batched forward run. Results equal ``FirstGradientGenerator`` per phantom.
    """

    def __init__(self, geometry, device, batch):
        from ..batch import BatchedAdjoint, BatchedForward

        self.base = FirstGradientGenerator(geometry, device)
        g, b = geometry, self.base
        self.g, self.device, self.P = geometry, device, int(batch)
        S, n, N = g.n_elements, g.n, g.n_steps
        shape = (S, n, n)
        e = lambda *sh: torch.empty(sh, device=device, dtype=_F32)  # noqa: E731
        self.truth = [TorchFDTD2D(TorchScalarWave2D(np.full((n, n), g.c0), g.spacing,
                                                    device=device), g.dt)
                      for _ in range(self.P)]
        self.obs_work = [[e(*shape) for _ in range(4)] for _ in range(self.P)]
        self.obs = [e(N, S, S) for _ in range(self.P)]
        rows = _i32(b.rows_np, device)
        cols = _i32(b.cols_np, device)
        values = _source_values_tensor(b.pulse, N, device)
        self.observe = BatchedForward([
            (self.truth[p], b.sampler, dict(n_shots=S, rows=rows, cols=cols, values=values,
                                            n_steps=N, init=(w[0], w[1]), work=w,
                                            samples=self.obs[p]))
            for p, w in enumerate(self.obs_work)]).prepare()
        self.residual = [e(N, S, S) for _ in range(self.P)]
        self.lam = [[e(*shape) for _ in range(4)] for _ in range(self.P)]
        self.grads = e(self.P, n, n)
        ws = int(_native.lib().swfd_adjoint_workspace_floats(n, n, S))
        self.ws = [e(ws) for _ in range(self.P)]
        eng = b.engine
        self.adjoint = []
        for j in range(eng.n_seg - 1, -1, -1):
            n0 = j * eng.K
            n1 = min(n0 + eng.K, N)
            self.adjoint.append(BatchedAdjoint([
                (b.solver0, b.sampler, dict(
                    n_shots=S, residual=self.residual[p], step_hi=n1, step_lo=n0,
                    state=(self.lam[p][0], self.lam[p][1]), work=self.lam[p], hist=eng.hist,
                    hist_first=n0 - 2, hist_every=1, image_scale=b.solver0.dt2,
                    grad=self.grads[p], workspace=self.ws[p]))
                for p in range(self.P)]).prepare())
        self.eye = torch.eye(S, device=device, dtype=torch.bool)

    def first_gradients(self, speeds):
        """Gradients (P, n, n) w.r.t. slowness squared for P speed maps."""
        from ..torch_backend import forward_run

        b, g = self.base, self.g
        for p, c in enumerate(speeds):
            self.truth[p].set_c(c)
            self.truth[p]._sync_medium()
            self.obs_work[p][0].zero_()
            self.obs_work[p][1].zero_()
        self.observe.launch()
        for p in range(self.P):
            r = torch.sub(b.samples0, self.obs[p], out=self.residual[p])
            r.masked_fill_(self.eye.unsqueeze(0), 0.0)  # transmitter's own channel muted
            self.lam[p][0].zero_()
            self.lam[p][1].zero_()
        self.grads.zero_()
        eng = b.engine
        for i, j in enumerate(range(eng.n_seg - 1, -1, -1)):
            n0 = j * eng.K
            n1 = min(n0 + eng.K, g.n_steps)
            if j == 0:
                eng.hist[0].zero_()
                eng.hist[1].zero_()
            else:
                eng.hist[0].copy_(eng.ckpt[2 * (j - 1)])
                eng.hist[1].copy_(eng.ckpt[2 * (j - 1) + 1])
            if n1 - 1 - n0 > 0:
                forward_run(b.solver0, eng.n_shots, eng.rows, eng.cols, eng.values, n1 - 1 - n0,
                            init=(eng.hist[0], eng.hist[1]), hist=eng.hist, hist_first=n0 - 2,
                            hist_every=1, step_begin=n0, work=eng.work, final=False)
            self.adjoint[i].launch()
        return -self.grads / (b.m0 * b.m0)


def target_from_speed(c, c0=1500.0, scale=100.0):
    return (c - c0) / scale


def speed_from_target(t, c0=1500.0, scale=100.0):
    return t * scale + c0


def generate(geometry, device, n_samples, seed0=0, log=None, log_every=30.0, batch=32):
    """Inputs (n_samples, n, n, 3), targets (n_samples, n, n) on the device, seconds.

    ``batch`` phantoms go through ``BatchedFirstGradients`` together (1: one
    at a time through ``FirstGradientGenerator``)."""
    n = geometry.n
    inputs = torch.empty((n_samples, n, n, 3), device=device, dtype=_F32)
    targets = torch.empty((n_samples, n, n), device=device, dtype=_F32)
    t0 = time.perf_counter()
    last = t0
    if batch > 1:
        bg = BatchedFirstGradients(geometry, device, batch)
        gen = bg.base
        for i0 in range(0, n_samples, batch):
            cs = [geometry.phantom(seed0 + i) for i in range(i0, min(n_samples, i0 + batch))]
            cs += [cs[-1]] * (batch - len(cs))  # pad the last batch
            grads = bg.first_gradients(cs)
            for k in range(min(batch, n_samples - i0)):
                inputs[i0 + k] = gen.make_input(grads[k])
                targets[i0 + k] = torch.as_tensor(target_from_speed(cs[k]), dtype=_F32,
                                                  device=device)
            now = time.perf_counter()
            if log is not None and now - last > log_every:
                torch.cuda.synchronize()
                log(f"generated {min(n_samples, i0 + batch)}/{n_samples} pairs in "
                    f"{now - t0:.1f} s")
                last = now
        torch.cuda.synchronize()
        return inputs, targets, time.perf_counter() - t0, gen
    gen = FirstGradientGenerator(geometry, device)
    for i in range(n_samples):
        c = geometry.phantom(seed0 + i)
        gm, _ = gen.first_gradient(c)
        inputs[i] = gen.make_input(gm)
        targets[i] = torch.as_tensor(target_from_speed(c), dtype=_F32, device=device)
        now = time.perf_counter()
        if log is not None and now - last > log_every:
            torch.cuda.synchronize()
            log(f"generated {i + 1}/{n_samples} pairs in {now - t0:.1f} s")
            last = now
    torch.cuda.synchronize()
    return inputs, targets, time.perf_counter() - t0, gen
