"""Multiscale FWI driver on the device objective.

A stage is one source wavelet (Ricker peak frequency or the reference chirp)
on one grid: observed data are simulated on the true model with that wavelet,
then an optimiser (Polak-Ribiere CG with Armijo line search by default, or
Adam / fixed-step GD) iterates ``TorchLeastSquaresFWI.misfit_and_grad``.
Stages run low to high frequency; the model of one stage starts the next,
moved between grids by :func:`resize_slowness` when the grid changes.

Restates the reference's stage runner (``invert_ring.py:1071-1250``), stage
ladder (``:1037-1045``) and bicubic stage transfer in physical coordinates
(``resize_slowness``, ``:1048-1068``). Plotting, CT media and argument parsing
are not part of this module.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from ..array import RingArray
from ..experiment import axis_centers, stable_dt
from ..sources import ricker
from ..torch_backend import TorchFDTD2D, TorchRingSampler, TorchScalarWave2D, simulate_shots_batch
from .fwi import TorchLeastSquaresFWI
from .medium import c_from_m, interior_mask, m_from_c
from .optim import adam, fixed_step_gradient_descent, polak_ribiere


def resize_slowness(m, source_axes, target_axes, background_m):
    """Bicubic transfer of slowness squared between cell-centred grids.

    Interpolates in physical coordinates (the two grids need not share a
    field of view); target points outside the source grid get
    ``background_m``. Host-side, float64.
    """
    sx, sy = (np.asarray(a, dtype=float) for a in source_axes)
    tx, ty = (np.asarray(a, dtype=float) for a in target_axes)
    u = 2.0 * (tx - sx[0]) / (sx[-1] - sx[0]) - 1.0
    v = 2.0 * (ty - sy[0]) / (sy[-1] - sy[0]) - 1.0
    gv, gu = np.meshgrid(v, u, indexing="ij")
    grid = torch.as_tensor(np.stack([gu, gv], axis=-1))[None]
    field_ = torch.as_tensor(np.asarray(m, dtype=float))[None, None]
    out = torch.nn.functional.grid_sample(field_, grid, mode="bicubic", padding_mode="border",
                                          align_corners=True)[0, 0].numpy()
    outside = (np.abs(gu) > 1.0) | (np.abs(gv) > 1.0)
    out[outside] = background_m
    return out


def rms_c_error(c_est, c_true, mask):
    err = np.asarray(c_est)[mask] - np.asarray(c_true)[mask]
    return float(np.sqrt(np.mean(err * err)))


@dataclass
class Stage:
    """One multiscale stage: grid, wavelet and iteration budget."""

    n: int
    spacing_m: float
    f_peak_hz: float
    n_steps: int
    iterations: int = 20


@dataclass
class StageResult:
    m: np.ndarray
    history: dict
    seconds: float
    rmse: list = field(default_factory=list)
    times: list = field(default_factory=list)
    stopped: bool = False  # the run reached stop_rmse in this stage


class _TargetReached(Exception):
    def __init__(self, m):
        super().__init__("target RMSE reached")
        self.m = m


def _objective_with_trace(problem, c_true_grid, mask, t0, rmse_out, times_out, cached=None,
                          stop_rmse=None, misfit_out=None):
    """Wrap misfit_and_grad to record interior RMS speed error and wall time.

    ``cached = (m, value, grad)`` answers the first call at exactly that ``m``;
    with ``stop_rmse`` the first evaluation at or below it ends the stage."""
    pending = [cached] if cached is not None else []

    def misfit_and_grad(m):
        if pending and np.array_equal(m, pending[0][0]):
            _, value, grad = pending.pop()
        else:
            pending.clear()
            value, grad = problem.misfit_and_grad(m)
        rmse_out.append(rms_c_error(c_from_m(problem.project(m)), c_true_grid, mask))
        times_out.append(time.perf_counter() - t0)
        if misfit_out is not None:
            misfit_out.append(float(value))
        if stop_rmse is not None and rmse_out[-1] <= stop_rmse:
            raise _TargetReached(m)
        return value, grad

    return misfit_and_grad


def run_stage(stage, c_true_fn, m_start, device, *, ring_radius, n_elements, c_min=1400.0,
              c_max=1650.0, c_background=1500.0, cfl=0.3, parallel_shots="auto",
              optimizer="cg", learning_rate=None, margin_cells=2, verbose=False, t0=None,
              start_hook=None, stop_rmse=None):
    """Run one stage; ``c_true_fn(x_m, y_m)`` gives the true speed on the stage grid.

    ``start_hook(problem, m_start, c_true, mask)`` may replace the start once the
    stage's objective exists (its time counts in the stage); it returns ``m`` or
    ``(m, (value, grad))`` when it has already evaluated the objective at ``m``.
    ``stop_rmse``: the stage ends at the first objective evaluation whose
    interior speed RMSE is at or below it (``StageResult.stopped``)."""
    t_start = time.perf_counter() if t0 is None else t0
    x = axis_centers(stage.n, stage.spacing_m)
    c_true = c_true_fn(x, x)
    dt = stable_dt(c_max, stage.spacing_m, stage.spacing_m, cfl=cfl)
    pulse = torch.as_tensor(ricker(dt, stage.f_peak_hz), dtype=torch.float32, device=device)
    array = RingArray(n_elements, ring_radius, x, x)
    model = TorchScalarWave2D(c_true, stage.spacing_m, device=device)
    solver = TorchFDTD2D(model, dt)
    sampler = TorchRingSampler(array, device=device)
    observed = torch.empty((n_elements, n_elements, stage.n_steps), device=device)
    per = 64
    for start in range(0, n_elements, per):
        sl = slice(start, min(n_elements, start + per))
        traces, _ = simulate_shots_batch(solver, sampler, pulse, array.rows[sl], array.cols[sl],
                                         stage.n_steps)
        observed[sl] = traces
    mask = interior_mask(x, x, ring_radius, margin_cells * stage.spacing_m)
    problem = TorchLeastSquaresFWI(solver, sampler, array, pulse, stage.n_steps, n_elements,
                                   observed, mask, c_min=c_min, c_max=c_max,
                                   c_background=c_background, parallel_shots=parallel_shots)
    cached = None
    if start_hook is not None:
        m_start = start_hook(problem, m_start, c_true, mask)
        if isinstance(m_start, tuple):
            m_start, (value, grad) = m_start
            m_start = problem.project(m_start)
            cached = (m_start, value, grad)
    rmse, times, misfits = [], [], []
    mg = _objective_with_trace(problem, c_true, mask, t_start, rmse, times, cached,
                               stop_rmse=stop_rmse, misfit_out=misfits)
    t_stage = time.perf_counter()
    try:
        m, history = _optimise(optimizer, m_start, problem, mg, mask, stage, learning_rate,
                               c_background, verbose)
    except _TargetReached as hit:
        return StageResult(m=problem.project(hit.m), history={"misfit": misfits},
                           seconds=time.perf_counter() - t_stage, rmse=rmse, times=times,
                           stopped=True), problem, c_true, mask
    return StageResult(m=m, history=history, seconds=time.perf_counter() - t_stage, rmse=rmse,
                       times=times), problem, c_true, mask


def _optimise(optimizer, m_start, problem, mg, mask, stage, learning_rate, c_background,
              verbose):
    if optimizer == "cg":
        m, history = polak_ribiere(m_start, problem.misfit, mg, problem.project, mask,
                                   max_iter=stage.iterations, verbose=verbose)
    elif optimizer == "adam":
        lr = learning_rate or 0.01 * float(np.mean(m_from_c(c_background)))
        m, history = adam(m_start, mg, problem.project, mask, lr, max_iter=stage.iterations,
                          verbose=verbose)
    elif optimizer == "gd":
        m, history = fixed_step_gradient_descent(m_start, mg, problem.project, mask,
                                                 learning_rate, max_iter=stage.iterations,
                                                 verbose=verbose)
    else:
        raise ValueError(f"Unknown optimizer {optimizer!r}; use cg, adam or gd.")
    return m, history


def multiscale_fwi(stages, c_true_fn, m0_fn, device, start_hook=None, on_stage=None,
                   **kwargs):
    """Run ``stages`` in order; ``m0_fn(x_m, y_m)`` gives the start on the first grid and
    ``start_hook`` (see :func:`run_stage`) may replace it inside the first stage;
    ``on_stage(results)`` is called after every stage."""
    t0 = time.perf_counter()
    results = []
    m = None
    axes = None
    background = m_from_c(kwargs.get("c_background", 1500.0))
    for stage in stages:
        x = axis_centers(stage.n, stage.spacing_m)
        if m is None:
            m_start = m0_fn(x, x)
        elif axes is not None and (len(axes[0]) != stage.n or axes[0][1] - axes[0][0] != x[1] - x[0]):
            m_start = resize_slowness(m, axes, (x, x), background)
        else:
            m_start = m
        result, problem, c_true, mask = run_stage(
            stage, c_true_fn, m_start, device, t0=t0,
            start_hook=start_hook if not results else None, **kwargs)
        results.append(result)
        if on_stage is not None:
            on_stage(results)
        if result.stopped:
            break
        m = result.m
        axes = (x, x)
    return results
