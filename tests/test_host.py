"""Host-side logic (CPU only): geometry, waveforms, regularisers, batching,
scatter table, phantoms and the benchmark's single-GPU contract."""

import os
import subprocess
import sys

import numpy as np
import pytest

from soundwavesimulation_amd import (ArcArray, RingArray, axis_centers, ellipse_phantom,
                                     linear_chirp, n_steps_for_crossing, ricker, source_values,
                                     stable_dt)
from soundwavesimulation_amd.inversion import (c_from_m, clip_m, interior_mask, m_from_c,
                                               tikhonov, total_variation, waveform_misfit)
from soundwavesimulation_amd.torch_backend import auto_batch_size, scatter_table

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
N, SPACING, RADIUS = 41, 2.0e-3, 0.030


def test_chirp_matches_reference_golden_pulse():
    g = np.load(os.path.join(GOLDEN, "ring41_fp64.npz"))
    pulse = linear_chirp(float(g["dt"]), 20.0e-6, 1.0e5, 2.5e5)
    np.testing.assert_allclose(pulse, g["pulse"], rtol=0, atol=1e-14)


def test_grid_helpers_closed_forms():
    np.testing.assert_allclose(axis_centers(5, 2.0), [-4.0, -2.0, 0.0, 2.0, 4.0])
    dt = stable_dt(1650.0, 0.5e-3, 0.5e-3, cfl=0.3)
    assert dt == pytest.approx(6.4282434653e-8, rel=1e-9)
    assert n_steps_for_crossing(0.03, 1500.0, 3.18e-7) == int(np.ceil(1.6 * 0.06 / 1500 / 3.18e-7))


def test_ricker_closed_form():
    f, dt = 300e3, 5e-9
    w = ricker(dt, f)
    delay = 1.5 / f
    k_peak = int(round(delay / dt))
    assert w[k_peak] == pytest.approx(1.0, abs=1e-12)
    assert w.size == int(np.ceil(2 * delay / dt)) + 1
    left = w[:k_peak + 1][::-1]
    np.testing.assert_allclose(w[k_peak:2 * k_peak + 1], left, atol=1e-12)  # even about the delay
    tz = delay + 1.0 / (np.sqrt(2.0) * np.pi * f)  # zero crossing
    assert ricker(dt, f, n_samples=1, delay_s=delay - tz)[0] == pytest.approx(0.0, abs=1e-12)
    assert abs(w).max() == pytest.approx(1.0)
    sv = source_values(w[:10], 20)
    assert sv.shape == (20,) and np.all(sv[10:] == 0) and np.all(sv[:10] == w[:10])


def test_ring_geometry_and_errors():
    x = axis_centers(N, SPACING)
    array = RingArray(6, RADIUS, x, x)
    assert array.inject_rows_cols(0) == (20, 35)
    with pytest.raises(ValueError, match="out of range"):
        array.inject_rows_cols(6)
    with pytest.raises(ValueError, match="too close to the x boundaries"):
        RingArray(6, 0.0395, x, x)
    with pytest.raises(ValueError, match="two elements would share a pixel"):
        RingArray(200, RADIUS, x, x)
    with pytest.raises(ValueError, match="Need at least 2"):
        RingArray(1, RADIUS, x, x)
    arc = ArcArray(8, RADIUS, 90.0, 45.0, x, x)
    assert arc.pitch_m == pytest.approx(RADIUS * np.pi / 2 / 7)
    ArcArray(64, RADIUS, 90.0, 45.0, x, x, allow_shared_pixels=True)
    with pytest.raises(ValueError, match="share a pixel"):
        ArcArray(64, RADIUS, 90.0, 45.0, x, x)


def test_scatter_table_matches_add_at_and_handles_collisions():
    x = axis_centers(N, SPACING)
    arc = ArcArray(40, RADIUS, 60.0, 10.0, x, x, allow_shared_pixels=True)
    r0, c0, w = arc.footprint()
    cell_map, start, cells, recs, ws = scatter_table(r0, c0, w, (N, N))
    rng = np.random.default_rng(0)
    res = rng.standard_normal(arc.n_elements)
    out = np.zeros(N * N)
    for e, cell in enumerate(cells):
        out[cell] = np.sum(res[recs[start[e]:start[e + 1]]] * ws[start[e]:start[e + 1]])
    np.testing.assert_allclose(out.reshape(N, N), arc.record_adjoint(res), rtol=1e-6, atol=1e-7)
    assert start[-1] == 4 * arc.n_elements
    assert np.any(np.diff(start) > 1)  # shared footprint cells exist
    assert np.all(cell_map[cells] == np.arange(cells.size))
    assert np.sum(cell_map >= 0) == cells.size


def test_medium_and_misfit():
    c = np.array([1400.0, 1500.0, 2000.0])
    np.testing.assert_allclose(c_from_m(m_from_c(c)), c)
    np.testing.assert_allclose(clip_m(m_from_c(np.array([1000.0, 3000.0])), 1400, 2000),
                               m_from_c(np.array([1400.0, 2000.0])))
    x = axis_centers(N, SPACING)
    mask = interior_mask(x, x, RADIUS, 2 * SPACING)
    assert mask[20, 20] and not mask[0, 0]
    with pytest.raises(ValueError):
        interior_mask(x, x, 0.001, 0.002)
    pred = np.ones((3, 5))
    obs = np.zeros((3, 5))
    assert waveform_misfit(pred, obs, 1) == pytest.approx(0.5 * 10)


def test_total_variation_gradient_and_scaling():
    rng = np.random.default_rng(1)
    x = axis_centers(N, SPACING)
    mask = interior_mask(x, x, RADIUS, margin_m=0.0)
    m = m_from_c(np.where(mask, 1500.0 + 100.0 * rng.standard_normal((N, N)), 1500.0))
    eps = 1e-9
    value, gradient = total_variation(m, mask, 1.0, eps)
    direction = np.where(mask, rng.standard_normal((N, N)), 0.0) * 1e-9
    step = 1e-3
    plus, _ = total_variation(m + step * direction, mask, 1.0, eps)
    minus, _ = total_variation(m - step * direction, mask, 1.0, eps)
    assert (plus - minus) / (2 * step) == pytest.approx(np.sum(gradient * direction), rel=1e-6)
    assert total_variation(np.full((N, N), 4e-7), mask, 1.0, eps)[0] == pytest.approx(0.0, abs=1e-20)
    assert np.all(gradient[~mask] == 0.0)
    assert value > 0.0


def test_tikhonov_gradient():
    rng = np.random.default_rng(2)
    x = axis_centers(N, SPACING)
    mask = interior_mask(x, x, RADIUS, 0.0)
    m = rng.standard_normal((N, N))
    val, g = tikhonov(m, mask, 3.0)
    d = np.where(mask, rng.standard_normal((N, N)), 0.0)
    h = 1e-6
    fd = (tikhonov(m + h * d, mask, 3.0)[0] - tikhonov(m - h * d, mask, 3.0)[0]) / (2 * h)
    assert fd == pytest.approx(np.sum(g * d), rel=1e-6)
    assert tikhonov(m, mask, 0.0) == (0.0, pytest.approx(np.zeros_like(m)))


def test_auto_batch_size_splits_evenly_on_cpu():
    assert auto_batch_size("cpu", 64, 1) == 16
    assert auto_batch_size("cpu", 20, 1) == 10
    assert auto_batch_size("cpu", 5, 1) == 5


def test_phantom_is_seeded():
    a = ellipse_phantom(64, 64, 0.5e-3, 4, seed=3, ring_radius_m=0.012)
    b = ellipse_phantom(64, 64, 0.5e-3, 4, seed=3, ring_radius_m=0.012)
    c = ellipse_phantom(64, 64, 0.5e-3, 4, seed=4, ring_radius_m=0.012)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 1400.0 and a.max() <= 1650.0


def test_bench_refuses_several_gpus():
    proc = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                          capture_output=True, text=True, timeout=120)
    assert proc.returncode == 2
    assert "one GPU" in proc.stderr


def test_warm_start_network_input():
    from soundwavesimulation_amd.surrogate.warm_start import block_mean, network_input

    rng = np.random.default_rng(0)
    a = rng.standard_normal((8, 8))
    assert np.allclose(block_mean(a, 2)[1, 2], a[2:4, 4:6].mean())
    assert np.allclose(block_mean(a, 8), a.mean())
    with pytest.raises(ValueError):
        block_mean(a, 3)
    mask = np.zeros((8, 8), dtype=bool)
    mask[2:6, 2:6] = True
    g = rng.standard_normal((8, 8))
    c = 1500.0 + 50.0 * rng.standard_normal((8, 8))
    x = network_input(g, c, mask, 2)
    assert x.shape == (4, 4, 4)
    assert np.isclose(np.max(np.abs(x[..., 0])), 1.0)
    assert np.all(x[0, :, 0] == 0.0)  # outside the mask: no gradient
    assert np.allclose(x[..., 1], (block_mean(c, 2) - 1500.0) / 100.0)
    assert np.allclose(x[0, :, 2], np.linspace(-1, 1, 4)) and np.allclose(x[:, 0, 3],
                                                                           np.linspace(-1, 1, 4))


def test_default_checkpoint_interval_is_near_sqrt_2n_and_a_multiple_of_four():
    import math

    from soundwavesimulation_amd.inversion.fwi import default_checkpoint_every, segment_plan

    for n in (20, 300, 1200, 1500, 4000, 5000, 6000, 20000):
        k = default_checkpoint_every(n)
        assert k % 4 == 0
        assert math.sqrt(2 * n) <= k < math.sqrt(2 * n) + 4
        n_seg, n_ckpt, n_hist = segment_plan(n, k)
        assert n_seg == -(-n // k) and n_ckpt == 2 * (n_seg - 1) and n_hist == k + 1
    assert default_checkpoint_every(6000) == 112
    assert default_checkpoint_every(5) == 3 and default_checkpoint_every(1) == 2


def test_auto_batch_size_never_admits_more_than_fits(monkeypatch):
    """On a device, the admitted batch times the per-shot bytes stays within the
    memory fraction of free + cached memory, and the split is even (the
    per-shot figure itself is checked against the measured peak on the GPU)."""
    import torch

    from soundwavesimulation_amd.inversion.fwi import gradient_bytes_per_shot

    free = 200 * 2**30
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda device=None: (free, 288 * 2**30))
    monkeypatch.setattr(torch.cuda, "memory_reserved", lambda device=None: 3 * 2**30)
    monkeypatch.setattr(torch.cuda, "memory_allocated", lambda device=None: 2**30)
    budget = 0.75 * (free + 2 * 2**30)
    cases = [(6000, (1024, 1024), 256, 1, None), (6000, (1024, 1024), 256, 4, torch.float16),
             (3039, (401, 401), 32, 4, torch.float16), (2490, (201, 201), 32, 1, None),
             (4000, (512, 512), 128, 2, torch.bfloat16), (5000, (768, 768), 128, 4, None)]
    for n_steps, shape, n_rec, every, dtype in cases:
        per_shot = gradient_bytes_per_shot(n_steps, shape, n_rec, every, dtype)
        for n_shots in (1, 7, 32, 128, 256, 1000):
            s = auto_batch_size("cuda", n_shots, per_shot)
            assert 1 <= s <= n_shots
            assert s == 1 or s * per_shot <= budget, (shape, every, dtype, n_shots, s)
            n_batches = -(-n_shots // s)
            assert n_batches * s - n_shots < n_batches  # an even split
    # reduced-precision histories are priced at their stored size (2 B per cell)
    full = gradient_bytes_per_shot(6000, (1024, 1024), 256, 4, torch.float32)
    half = gradient_bytes_per_shot(6000, (1024, 1024), 256, 4, torch.float16)
    assert full - half == 1500 * 1024 * 1024 * 2


def test_wave_plan_takes_shortest_work_first_within_memory_and_width():
    from soundwavesimulation_amd.batch import plan_waves
    from soundwavesimulation_amd.uneven import draw, sizes_of

    sizes = sizes_of(draw(2))
    budget, width = 60e9, 48e6
    waves = plan_waves(sizes, budget, width)
    flat = [i for w in waves for i in w]
    assert sorted(flat) == list(range(24))
    work = [sizes[i][0] for i in flat]
    assert work == sorted(work)  # shortest work first
    for w in waves:
        assert sum(sizes[i][1] for i in w) <= budget
        assert len(w) == 1 or sum(sizes[i][2] for i in w) <= width
    # any draw runs within a budget that holds its largest problem
    big = max(z[1] for z in sizes)
    assert all(sum(sizes[i][1] for i in w) <= big for w in plan_waves(sizes, big, 1e12))
    with pytest.raises(ValueError, match="more than the budget"):
        plan_waves(sizes, 0.5 * big, 1e12)
