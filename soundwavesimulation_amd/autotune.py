"""Launch-shape autotuner for the stencil kernels (one grid size at a time).

For a grid (ny, nx) and a shot count it times, on the device, short runs of
the forward and adjoint kernels under candidate shapes and installs the
fastest through ``swfd_tune`` (``include/swfd.h``):

* single-step kernels for both (kind 2, depth 1);
* forward: two-step marches in 64-column blocks (rows per march x target
  work items per CU), or full-width workgroups with 2-4 steps per launch
  (rows per march);
* adjoint: two-step blocks (rows per march <= 32 x shot groups), or
  full-width workgroups with 2-4 steps per launch (rows per march; the
  segment's image lives in LDS; shot groups chosen to fill the CUs).

The batched lockstep launches use the same forms. Results are cached per
device name, grid and shot count in a JSON file (``SWFD_TUNE_CACHE``, default
``~/.cache/swfd_tune.json``), so a later process applies them without timing
again.
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
WIDE_LEVELS = (2, 3, 4)
WIDE_FORWARD_SEG_ROWS = (16, 32, 64, 128)
WIDE_ADJOINT_SEG_ROWS = (8, 16, 32)


def _cache_path():
    return os.environ.get("SWFD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "swfd_tune.json"))


def _key(device, ny, nx, n_shots, adjoint=True):
    return f"{torch.cuda.get_device_name(device)}:{ny}x{nx}:{n_shots}:v2" + ("" if adjoint else ":fwd")


def apply(cfg, ny, nx):
    """Install a configuration returned by ``tune``."""
    lib = _native.lib()
    f, a = cfg["forward"], cfg["adjoint"]
    _native.check(lib.swfd_tune(0, ny, nx, f["seg"], f.get("wpc", 0), 0))
    if a is not None:
        _native.check(lib.swfd_tune(1, ny, nx, a["seg"], a.get("groups", 0), 0))
    if cfg["depth1"]:
        _native.check(lib.swfd_tune(2, ny, nx, 0, 0, 1))
        return
    _native.check(lib.swfd_tune(2, ny, nx, 0, f["wide"], f["levels"]))
    if a is not None:
        _native.check(lib.swfd_tune(3, ny, nx, 0, a["wide"], a["levels"]))


def _time(fn, reps=2, label=None, log=None):
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    t = a.elapsed_time(b) / 1e3 / reps
    if log is not None and label is not None:
        log(f"    {label}: {t * 1e3:.3f} ms")
    return t


def tune(ny, nx, n_shots, device, n_steps=16, spacing=0.5e-3, use_cache=True, log=None,
         verbose=None, adjoint=True):
    """Time candidates for (ny, nx, n_shots); install and return the fastest
    (``verbose``: a log function for every candidate's time). ``adjoint=False``
    tunes the forward launches only and leaves the adjoint's shape to the
    library's model (the shapes here are timed on the adjoint over a stored
    field history, not on the adjoint that rebuilds the forward fields)."""
    path = _cache_path()
    key = _key(device, ny, nx, n_shots, adjoint)
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
    n_steps = max(12, n_steps // 12 * 12)  # whole launches for 1-4 steps per launch
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

    # entries of this grid only (other grids keep theirs); depth 0 = no entry
    for kind in (0, 1, 2, 3):
        lib.swfd_tune(kind, ny, nx, 0, 0, 0)
    wide = bool(lib.swfd_lap_history(ny, nx))  # full-width launches fit this grid
    lib.swfd_tune(2, ny, nx, 0, 0, 1)
    t1 = _time(fwd) + _time(adj)
    # forward candidates
    fwd_best = None
    lib.swfd_tune(2, ny, nx, 0, 0, 2)
    for seg in FORWARD_SEG_ROWS:
        if seg > ny:
            continue
        for wpc in FORWARD_WAVES_PER_CU:
            lib.swfd_tune(0, ny, nx, seg, wpc, 0)
            t = _time(fwd, label=f"forward blocks seg={seg} wpc={wpc}", log=verbose)
            if fwd_best is None or t < fwd_best[0]:
                fwd_best = (t, {"wide": 0, "levels": 2, "seg": seg, "wpc": wpc})
    if wide:
        for levels in WIDE_LEVELS:
            lib.swfd_tune(2, ny, nx, 0, 1, levels)
            for seg in WIDE_FORWARD_SEG_ROWS:
                if seg > ny:
                    continue
                lib.swfd_tune(0, ny, nx, seg, 0, 0)
                t = _time(fwd, label=f"forward wide L={levels} seg={seg}", log=verbose)
                if t < fwd_best[0]:
                    fwd_best = (t, {"wide": 1, "levels": levels, "seg": seg, "wpc": 0})
    # adjoint candidates
    adj_best = None
    lib.swfd_tune(2, ny, nx, 0, 0, 2)
    lib.swfd_tune(3, ny, nx, 0, 0, 2)
    for seg in (ADJOINT_SEG_ROWS if adjoint else ()):
        if seg > ny:
            continue
        for groups in ADJOINT_GROUPS:
            lib.swfd_tune(1, ny, nx, seg, groups, 0)
            t = _time(adj, label=f"adjoint blocks seg={seg} groups={groups}", log=verbose)
            if adj_best is None or t < adj_best[0]:
                adj_best = (t, {"wide": 0, "levels": 2, "seg": seg, "groups": groups})
    if wide and adjoint:
        for levels in WIDE_LEVELS:
            lib.swfd_tune(3, ny, nx, 0, 1, levels)
            for seg in WIDE_ADJOINT_SEG_ROWS:
                if seg > ny:
                    continue
                lib.swfd_tune(1, ny, nx, seg, 0, 0)  # shot groups: automatic
                t = _time(adj, label=f"adjoint wide L={levels} seg={seg}", log=verbose)
                if t < adj_best[0]:
                    adj_best = (t, {"wide": 1, "levels": levels, "seg": seg, "groups": 0})
    if log is not None:
        log(f"  {ny}x{nx}: single {t1 / n_steps * 1e3:.4f}, forward {fwd_best}, adjoint {adj_best}"
            " (ms per step for the whole run)")
    if adj_best is None:  # forward only: the adjoint keeps the library's shape
        for kind in (1, 3):
            lib.swfd_tune(kind, ny, nx, 0, 0, 0)
        lib.swfd_tune(2, ny, nx, 0, 0, 1)
        t1 = _time(fwd)
        adj_best = (0.0, None)
    t2 = fwd_best[0] + adj_best[0]
    cfg = {"depth1": t1 < t2, "forward": fwd_best[1], "adjoint": adj_best[1],
           "ms_per_step_single": round(t1 / n_steps * 1e3, 4),
           "ms_per_step_multi": round(t2 / n_steps * 1e3, 4)}
    apply(cfg, ny, nx)
    if log is not None:
        log(f"tuned {ny}x{nx} x {S} shots: {cfg}")
    if use_cache:
        cache[key] = cfg
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1)
    return cfg
