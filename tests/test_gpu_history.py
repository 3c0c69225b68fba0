"""Reduced-precision wavefield histories held in float16 / bfloat16 on the device
(the reference's ``store_every`` + ``wave_dtype`` path, ``invert_ring.py:422-455,
683-701``):

* ``swfd_history_convert`` rounds to nearest even exactly as torch's own casts do
  (bit for bit, subnormals, overflow to inf, signed zeros, ragged lengths and
  unaligned slices taking the one-value-per-lane kernel);
* ``forward_store_batch`` returns the history in ``wave_dtype``, equal bit for
  bit to the float32 history cast by torch, with the same traces;
* the chunked adjoint over the packed history equals the adjoint over the
  float32 history rounded through the same dtype (rel-L2 <= 1e-6: only the
  order of the per-chunk image sums differs);
* the device memory a reduced-precision or exact gradient batch holds stays
  within ``gradient_bytes_per_shot`` per shot, so ``parallel_shots="auto"``
  never admits more shots than fit.
"""

import os

import numpy as np
import pytest
import torch

from soundwavesimulation_amd import RingArray, axis_centers, ricker, stable_dt
from soundwavesimulation_amd.inversion import interior_mask
from soundwavesimulation_amd.inversion.fwi import (TorchLeastSquaresFWI, forward_store_batch,
                                                   gradient_bytes_per_shot)
from soundwavesimulation_amd.phantoms import ellipse_phantom
from soundwavesimulation_amd.torch_backend import (TorchFDTD2D, TorchFDTD2DBatch,
                                                   TorchRingSampler, TorchScalarWave2D,
                                                   _source_values_tensor, convert_history,
                                                   simulate_shots_batch)

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DEV = torch.device("cuda", 0)
HALF = [torch.float16, torch.bfloat16]


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def bits(t):
    return t.view(torch.int16).cpu()


@pytest.mark.parametrize("dtype", HALF)
@pytest.mark.parametrize("n, offset", [(8 * 4096, 0), (1001, 0), (8 * 512, 3)])
def test_history_convert_rounds_like_torch(dtype, n, offset):
    g = torch.Generator().manual_seed(n + offset)
    mag = torch.pow(10.0, torch.empty(n + offset).uniform_(-12.0, 6.0, generator=g))
    x = (torch.randn(n + offset, generator=g).sign() * mag).float()
    special = torch.tensor([0.0, -0.0, 65504.0, 65520.0, -1e6, 6.0e-8, 2.9e-8, 5.96e-8,
                            1e-40, 3.0e38, 1.0 + 2.0**-11, 1.0 + 3 * 2.0**-11,
                            1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8])
    x[offset:offset + special.numel()] = special
    src = x.to(DEV)[offset:]
    packed = torch.empty(n, device=DEV, dtype=dtype)
    convert_history(src, packed)
    expect = src.to(dtype)
    assert torch.equal(bits(packed), bits(expect))
    back = torch.empty(n, device=DEV, dtype=torch.float32)
    convert_history(packed, back)
    assert torch.equal(back.cpu(), expect.float().cpu())


def ring41():
    g = dict(np.load(os.path.join(GOLDEN, "ring41_fp32.npz")))
    sp, dt = float(g["spacing"]), float(g["dt"])
    x = axis_centers(41, sp)
    array = RingArray(6, float(g["radius"]), x, x)
    solver = TorchFDTD2D(TorchScalarWave2D(g["c_true"], sp, device=DEV), dt)
    sampler = TorchRingSampler(array, device=DEV)
    return g, array, solver, sampler


@pytest.mark.parametrize("dtype", HALF)
@pytest.mark.parametrize("store_every", [1, 4])
def test_forward_store_batch_holds_the_history_in_wave_dtype(dtype, store_every):
    g, array, solver, sampler = ring41()
    N = int(g["n_steps"])
    values = _source_values_tensor(torch.as_tensor(g["pulse"], dtype=torch.float32), N, DEV)
    sources = np.arange(6)
    b32 = TorchFDTD2DBatch(solver, 6)
    t32, w32 = forward_store_batch(solver, sampler, array, values, sources, N, b32,
                                   store_every, torch.float32)
    b16 = TorchFDTD2DBatch(solver, 6)
    t16, w16 = forward_store_batch(solver, sampler, array, values, sources, N, b16,
                                   store_every, dtype)
    assert w16.dtype == dtype and w16.shape == w32.shape
    assert torch.equal(t16, t32)
    assert torch.equal(bits(w16), bits(w32.to(dtype)))
    assert torch.equal(b16.u, b32.u) and torch.equal(b16.u_old, b32.u_old)
    traces, _ = simulate_shots_batch(solver, sampler, values, array.rows, array.cols, N)
    assert torch.equal(t16, traces)


@pytest.mark.parametrize("dtype", HALF)
@pytest.mark.parametrize("store_every", [1, 2, 4])
def test_packed_history_adjoint_equals_rounded_float32_history(dtype, store_every):
    g, array, solver, sampler = ring41()
    N = int(g["n_steps"])
    problem = TorchLeastSquaresFWI(solver, sampler, array, g["pulse"], N, 6,
                                   torch.as_tensor(g["observed"]), g["mask"], c_min=1400.0,
                                   c_max=2000.0, store_every=store_every, wave_dtype=dtype)
    problem.apply_medium(g["m0"])
    b = TorchFDTD2DBatch(solver, 6)
    sources = np.arange(6)
    traces, w16 = forward_store_batch(solver, sampler, array, problem.source_values, sources,
                                      N, b, store_every, dtype)
    _, w32 = forward_store_batch(solver, sampler, array, problem.source_values, sources, N,
                                 TorchFDTD2DBatch(solver, 6), store_every, torch.float32)
    residual = problem._residual_batch(traces, problem.observed[:6], sources)
    g_packed = torch.zeros((41, 41), device=DEV)
    problem._adjoint_accumulate_batch(b, w16, residual, g_packed)
    g_rounded = torch.zeros((41, 41), device=DEV)
    problem._adjoint_accumulate_batch(TorchFDTD2DBatch(solver, 6), w32.to(dtype).float(),
                                      residual, g_rounded)
    e = rel(g_packed.cpu().numpy(), g_rounded.cpu().numpy())
    print(f"{dtype} store_every={store_every}: packed vs rounded float32 history {e:.2e}")
    assert e <= 1e-6


def _peak_bytes(fn):
    torch.cuda.synchronize(DEV)
    torch.cuda.empty_cache()
    base = torch.cuda.memory_allocated(DEV)
    torch.cuda.reset_peak_memory_stats(DEV)
    fn()
    torch.cuda.synchronize(DEV)
    return torch.cuda.max_memory_allocated(DEV) - base


@pytest.mark.parametrize("options", [dict(store_every=4, wave_dtype=torch.float16),
                                     dict(store_every=2, wave_dtype=torch.bfloat16),
                                     dict(store_every=4, wave_dtype=torch.float32),
                                     dict()])
def test_gradient_batch_memory_stays_within_the_per_shot_model(options):
    n, h, radius, n_el, N, S = 192, 0.5e-3, 0.04, 32, 1200, 16
    c_true = ellipse_phantom(n, n, h, 3, seed=4, ring_radius_m=radius)
    x = axis_centers(n, h)
    dt = stable_dt(1650.0, h, h, cfl=0.3)
    pulse = ricker(dt, 300.0e3)
    array = RingArray(n_el, radius, x, x)
    truth = TorchFDTD2D(TorchScalarWave2D(c_true, h, device=DEV), dt)
    sampler = TorchRingSampler(array, device=DEV)
    observed, _ = simulate_shots_batch(truth, sampler,
                                       torch.as_tensor(pulse, dtype=torch.float32),
                                       array.rows, array.cols, N)
    solver = TorchFDTD2D(TorchScalarWave2D(np.full((n, n), 1500.0), h, device=DEV), dt)
    mask = interior_mask(x, x, radius, 2 * h)
    problem = TorchLeastSquaresFWI(solver, sampler, array, pulse, N, S, observed, mask,
                                   parallel_shots=S, **options)
    m0 = np.full((n, n), 1.0 / 1500.0**2)
    # the first gradient allocates every buffer the batch holds (engines included)
    peak = _peak_bytes(lambda: problem.misfit_and_grad(m0))
    per_shot = gradient_bytes_per_shot(N, (n, n), n_el, options.get("store_every", 1),
                                       options.get("wave_dtype"), torch.float32,
                                       problem.checkpoint_every)
    print(f"{options}: peak {peak / 2**20:.1f} MiB for {S} shots, model "
          f"{S * per_shot / 2**20:.1f} MiB")
    assert peak <= S * per_shot
