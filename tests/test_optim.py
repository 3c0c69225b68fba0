"""Optimisers and stage transfer (CPU only, on analytic objectives)."""

import numpy as np
import pytest

from soundwavesimulation_amd import axis_centers
from soundwavesimulation_amd.inversion import m_from_c
from soundwavesimulation_amd.inversion.multiscale import resize_slowness, rms_c_error
from soundwavesimulation_amd.inversion.optim import (adam, fixed_step_gradient_descent,
                                                     masked_dot, polak_ribiere)


class Quadratic:
    """J(m) = 1/2 (m - t)^T A (m - t) on the mask, A diagonal positive."""

    def __init__(self, shape=(12, 12), seed=0):
        rng = np.random.default_rng(seed)
        self.mask = np.zeros(shape, dtype=bool)
        self.mask[2:-2, 2:-2] = True
        self.a = 1.0 + 4.0 * rng.random(shape)
        self.t = rng.standard_normal(shape)
        self.calls = 0

    def misfit(self, m):
        d = np.where(self.mask, m - self.t, 0.0)
        return 0.5 * float(np.sum(self.a * d * d))

    def misfit_and_grad(self, m):
        self.calls += 1
        d = np.where(self.mask, m - self.t, 0.0)
        return 0.5 * float(np.sum(self.a * d * d)), self.a * d

    def project(self, m):
        out = np.array(m, dtype=float, copy=True)
        out[~self.mask] = 0.0
        return out


def test_polak_ribiere_converges_on_quadratic():
    q = Quadratic()
    # the first trial step is a fraction of ||m|| (slowness squared is never 0)
    m0 = q.project(np.full(q.mask.shape, 0.5))
    m, hist = polak_ribiere(m0, q.misfit, q.misfit_and_grad, q.project, q.mask, max_iter=60,
                            verbose=False)
    assert hist["misfit"][-1] < 1e-6 * hist["misfit"][0]
    assert all(b <= a + 1e-12 for a, b in zip(hist["misfit"], hist["misfit"][1:]))
    np.testing.assert_allclose(m[q.mask], q.t[q.mask], atol=1e-3)
    assert len(hist["step_size"]) == len(hist["misfit"]) == len(hist["grad_norm"])


def test_gd_and_adam_decrease_misfit():
    q = Quadratic(seed=1)
    m0 = np.zeros(q.mask.shape)
    _, h_gd = fixed_step_gradient_descent(m0, q.misfit_and_grad, q.project, q.mask, 0.15,
                                          max_iter=30, verbose=False)
    assert h_gd["misfit"][-1] < 1e-3 * h_gd["misfit"][0]
    _, h_adam = adam(m0, q.misfit_and_grad, q.project, q.mask, 0.05, max_iter=200, verbose=False)
    assert h_adam["misfit"][-1] < 1e-2 * h_adam["misfit"][0]
    assert masked_dot(np.ones((2, 2)), np.full((2, 2), 3.0), np.eye(2, dtype=bool)) == 6.0


def test_line_search_failure_stops():
    q = Quadratic(seed=2)
    bad_misfit = lambda m: float("inf")  # noqa: E731
    m, hist = polak_ribiere(np.zeros(q.mask.shape), bad_misfit, q.misfit_and_grad, q.project,
                            q.mask, max_iter=5, verbose=False)
    assert len(hist["misfit"]) == 1 and np.all(m == 0)


def test_stage_transfer_uses_physical_coordinates():
    def disk_model(x_m, y_m):
        x, y = np.meshgrid(x_m, y_m, indexing="xy")
        return m_from_c(np.where(x**2 + (y - 0.03)**2 <= 0.02**2, 1600.0, 1480.0))

    source = (axis_centers(201, 1.828e-3), axis_centers(201, 1.828e-3))
    target = (axis_centers(401, 1.5e-3), axis_centers(401, 1.5e-3))
    moved = resize_slowness(disk_model(*source), source, target, m_from_c(1480.0))
    expected = disk_model(*target)
    inside = np.abs(expected - moved) > 0.5 * abs(m_from_c(1600.0) - m_from_c(1480.0))
    assert inside.mean() < 0.002
    assert moved[0, 0] == pytest.approx(m_from_c(1480.0))
    assert rms_c_error(np.full((3, 3), 1500.0), np.full((3, 3), 1490.0),
                       np.ones((3, 3), bool)) == pytest.approx(10.0)
