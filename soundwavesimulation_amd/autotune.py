"""Launch-shape autotuner for the stencil kernels (one grid size at a time).

For a grid (ny, nx) and a shot count it times, on the device, short runs of
the forward and adjoint kernels under candidate shapes and installs the
fastest through ``swfd_tune`` (``include/swfd.h``):

* temporal depth: single-step kernels (depth 1) or two-step marches (depth 2);
* forward marches: rows per march (the tile is 64 columns by that many rows)
  and the target number of work items per CU (which sets the shots per
  wavefront);
* adjoint marches: rows per march (<= 32, bounded by the imaging accumulator
  in LDS) and the number of shot groups (<= 8 partial images).

Results are cached per device name, grid and shot count in a JSON file
(``SWFD_TUNE_CACHE``, default ``~/.cache/swfd_tune.json``), so a later
process applies them without timing again.
"""

from __future__ import annotations

import json
import os

import numpy as np
import torch

from . import _native
from .array import RingArray
from .experiment import axis_centers, stable_dt
from .torch_backend import (TorchFDTD2D, TorchRingSampler, TorchScalarWave2D, adjoint_run,
                            adjoint_workspace, forward_run)

FORWARD_SEG_ROWS = (8, 16, 32, 64, 128)
FORWARD_WAVES_PER_CU = (128, 256)
ADJOINT_SEG_ROWS = (8, 16, 32)
ADJOINT_GROUPS = (4, 8)


def _cache_path():
    return os.environ.get("SWFD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "swfd_tune.json"))


def _key(device, ny, nx, n_shots):
    return f"{torch.cuda.get_device_name(device)}:{ny}x{nx}:{n_shots}"


def apply(cfg, ny, nx):
    lib = _native.lib()
    _native.check(lib.swfd_tune(0, ny, nx, cfg["fwd_seg"], cfg["fwd_wpc"], 0))
    _native.check(lib.swfd_tune(1, ny, nx, cfg["adj_seg"], cfg["adj_groups"], 0))
    _native.check(lib.swfd_tune(2, ny, nx, 0, 0, cfg["depth"]))


def _time(fn, reps=2):
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / 1e3 / reps


def tune(ny, nx, n_shots, device, n_steps=16, spacing=0.5e-3, use_cache=True, log=None):
    """Time candidates for (ny, nx, n_shots); install and return the fastest."""
    path = _cache_path()
    key = _key(device, ny, nx, n_shots)
    cache = {}
    if use_cache and os.path.exists(path):
        with open(path) as f:
            cache = json.load(f)
        if key in cache:
            apply(cache[key], ny, nx)
            return cache[key]
    lib = _native.lib()
    dt = stable_dt(1650.0, spacing, spacing, cfl=0.3)
    solver = TorchFDTD2D(TorchScalarWave2D(np.full((ny, nx), 1500.0), spacing, device=device), dt)
    x, y = axis_centers(nx, spacing), axis_centers(ny, spacing)
    n_el = max(2, min(n_shots, 256))
    array = RingArray(n_el, 0.4 * min(ny, nx) * spacing, x, y)
    sampler = TorchRingSampler(array, device=device)
    S = n_shots
    rows = torch.as_tensor(np.resize(array.rows, S).astype(np.int32), device=device)
    cols = torch.as_tensor(np.resize(array.cols, S).astype(np.int32), device=device)
    vals = torch.zeros(n_steps + 2, device=device)
    vals[:4] = 1.0
    shape = (S, ny, nx)
    work = [torch.zeros(shape, device=device) for _ in range(4)]
    hist = torch.zeros((n_steps + 2, *shape), device=device)
    res = torch.zeros((n_steps, S, n_el), device=device)
    grad = torch.zeros((ny, nx), device=device)
    ws = adjoint_workspace(solver, S)

    def fwd():
        forward_run(solver, S, rows, cols, vals, n_steps, work=work, final=False)

    def adj():
        adjoint_run(solver, sampler, S, res, n_steps, 0, state=(work[0], work[1]), work=work,
                    hist=hist, hist_first=-2, hist_every=1, image_scale=1e-15, grad=grad,
                    workspace=ws)

    results = []
    lib.swfd_tune(2, ny, nx, 0, 0, 1)
    t1 = _time(fwd) + _time(adj)
    results.append(("depth1", t1))
    lib.swfd_tune(2, ny, nx, 0, 0, 2)
    best_f = None
    for seg in FORWARD_SEG_ROWS:
        if seg > ny:
            continue
        for wpc in FORWARD_WAVES_PER_CU:
            lib.swfd_tune(0, ny, nx, seg, wpc, 0)
            t = _time(fwd)
            if best_f is None or t < best_f[0]:
                best_f = (t, seg, wpc)
    lib.swfd_tune(0, ny, nx, best_f[1], best_f[2], 0)
    best_a = None
    for seg in ADJOINT_SEG_ROWS:
        if seg > ny:
            continue
        for groups in ADJOINT_GROUPS:
            lib.swfd_tune(1, ny, nx, seg, groups, 0)
            t = _time(adj)
            if best_a is None or t < best_a[0]:
                best_a = (t, seg, groups)
    t2 = best_f[0] + best_a[0]
    cfg = {"depth": 2 if t2 <= t1 else 1, "fwd_seg": best_f[1], "fwd_wpc": best_f[2],
           "adj_seg": best_a[1], "adj_groups": best_a[2],
           "ms_per_step_depth1": round(t1 / n_steps * 1e3, 4),
           "ms_per_step_depth2": round(t2 / n_steps * 1e3, 4)}
    apply(cfg, ny, nx)
    if log is not None:
        log(f"tuned {ny}x{nx} x {S} shots: {cfg}")
    if use_cache:
        cache[key] = cfg
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1)
    return cfg
