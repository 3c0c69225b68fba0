"""Benchmark: one full-waveform-inversion gradient on one MI355X.

Workload (default): 1024 x 1024 grid at 0.25 mm, 256-element ring of radius
110 mm, all 256 shots per gradient, 6000 steps, 300 kHz Ricker, true medium
of 12 seeded ellipses (c 1400-1650 m/s), start = truth blurred by 10 mm.
A "step" is one call of ``TorchLeastSquaresFWI.misfit_and_grad`` (forward
with checkpoints and the boundary ring, residual and misfit, adjoint with
imaging that rebuilds the forward field backwards in time -- or, with
SWFD_RECON=0, segment recomputation and an adjoint over the recomputed
fields --, host chain rule to slowness squared). Inputs are synthetic and
resident in HBM before the timed region.

Printed JSON (one line): Gcell/s = 2 * ny * nx * steps * shots / seconds per
gradient, plus
  roofline      -- the fused adjoint kernel (L steps per launch, L from the
                   library): `achieved` = the launch's algorithmic bytes (its
                   inputs read once and outputs written once) over its average
                   launch time, timed with HIP events over at least 50
                   launches on the stream they run on. Reconstructing
                   (adjr_kernel): per shot the adjoint pair and the forward
                   pair read and written, the L steps' boundary cells; c^2,
                   the residual rows and the gradient. Recomputing
                   (adjm_kernel): lam_old, lam and the L forward fields of
                   every shot read, lam of the last two levels written, c^2,
                   the residual rows and the gradient. `traffic` = the HBM
                   bytes per launch that rocprofv3's FETCH_SIZE / WRITE_SIZE
                   counters measured for the same launch shape
                   (profiles/traffic_adjoint[_recon]_<grid>.json, keyed on
                   kernel, grid, steps, shots, segment rows and shot groups;
                   null with the reason when the shape differs),
                   `traffic_frac` the rate those bytes imply;
                   `algorithmic_equiv` restates the launch at 20 B per
                   cell-update (a single step's streams), a figure that
                   multi-step launches exceed;
  cpu_baseline  -- the float64 C oracle (OpenMP, all host cores) on a
                   bounded sample of the same workload, with the CPU model;
  torch_reference(_compiled) -- the reference's op sequence (conv2d
                   Laplacian, elementwise leapfrog, index_put_ sources, Mur
                   slices, bilinear gather, imaging) in PyTorch on the same
                   GPU, eager and under torch.compile(mode="reduce-overhead",
                   fullgraph=True) as the reference compiles it, timed over a
                   few steps and scaled to one gradient.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2, help="timed gradients")
    p.add_argument("--warmup", type=int, default=1, help="untimed gradients")
    p.add_argument("--grid", type=int, default=1024)
    p.add_argument("--spacing-mm", type=float, default=0.25)
    p.add_argument("--shots", type=int, default=256)
    p.add_argument("--nsteps", type=int, default=6000)
    p.add_argument("--radius-mm", type=float, default=None,
                   help="ring radius (default 110 mm on the 256 mm grid, same fraction otherwise)")
    p.add_argument("--freq-khz", type=float, default=300.0)
    p.add_argument("--checkpoint-every", type=int, default=0)
    p.add_argument("--parallel-shots", default="auto")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-torch-reference", action="store_true")
    p.add_argument("--cpu-shots", type=int, default=4)
    p.add_argument("--cpu-steps", type=int, default=1000)
    p.add_argument("--copy-calibration", action="store_true",
                   help="run only a plain streaming kernel of known size (c = a + b over two "
                        "2 GiB float32 inputs, 5 launches), to scale HBM traffic counters")
    p.add_argument("--dry-run", action="store_true",
                   help="build the host-side geometry only and print it (no GPU)")
    a = p.parse_args()
    if a.radius_mm is None:
        a.radius_mm = 110.0 * (a.grid * a.spacing_mm) / 256.0
    return a


def host_geometry(a):
    from soundwavesimulation_amd import RingArray, axis_centers, ricker, stable_dt
    from soundwavesimulation_amd.phantoms import ellipse_phantom

    n = a.grid
    h = a.spacing_mm * 1e-3
    radius = a.radius_mm * 1e-3
    c_true = ellipse_phantom(n, n, h, 12, seed=5, ring_radius_m=radius)
    x = axis_centers(n, h)
    dt = stable_dt(1650.0, h, h, cfl=0.3)
    pulse = ricker(dt, a.freq_khz * 1e3)
    array = RingArray(a.shots, radius, x, x)
    return c_true, x, dt, h, radius, pulse, array


def build_problem(a, dev):
    from soundwavesimulation_amd.inversion import interior_mask
    from soundwavesimulation_amd.inversion.fwi import TorchLeastSquaresFWI
    from soundwavesimulation_amd.phantoms import blurred_start
    from soundwavesimulation_amd.torch_backend import (TorchFDTD2D, TorchRingSampler,
                                                       TorchScalarWave2D, simulate_shots_batch)

    c_true, x, dt, h, radius, pulse, array = host_geometry(a)
    model = TorchScalarWave2D(c_true, h, device=dev)
    solver = TorchFDTD2D(model, dt)
    sampler = TorchRingSampler(array, device=dev)
    observed = torch.empty((a.shots, a.shots, a.nsteps), device=dev)
    pulse_t = torch.as_tensor(pulse, dtype=torch.float32, device=dev)
    per = 64
    for start in range(0, a.shots, per):
        sl = slice(start, min(a.shots, start + per))
        tr, _ = simulate_shots_batch(solver, sampler, pulse_t, array.rows[sl], array.cols[sl],
                                     a.nsteps)
        observed[sl] = tr
        del tr
    torch.cuda.synchronize()
    mask = interior_mask(x, x, radius, 2 * h)
    m0 = blurred_start(c_true, 10e-3, h, mask)
    ps = a.parallel_shots if a.parallel_shots == "auto" else int(a.parallel_shots)
    problem = TorchLeastSquaresFWI(solver, sampler, array, pulse_t, a.nsteps, a.shots, observed,
                                   mask, c_min=1400.0, c_max=1650.0, c_background=1500.0,
                                   parallel_shots=ps,
                                   checkpoint_every=a.checkpoint_every or None)
    torch.cuda.empty_cache()
    return problem, m0, c_true, dt, h


def time_adjoint_kernel(problem, min_launches=50):
    """Average launch time of the gradient's adjoint kernel (residual
    injection + imaging, and with reconstruction the backward rebuild of the
    forward field).

    A run advances ``L`` adjoint steps per launch (``swfd_steps_per_launch``).
    Recomputing gradients: runs of the segment length (the history holds that
    many fields) from step 0; reconstructing gradients: runs of 16 launches
    over the last steps of the record (every level images), rebuilding from
    the forward run's final pair and ring. Runs repeat until at least
    ``min_launches`` launches are timed, with HIP events on torch's current
    stream, where the library launches. Each run also clears and sums the shot
    groups' partial images once (two small kernels, about 0.2 % of the timed
    region here).
    """
    import soundwavesimulation_amd._native as native
    from soundwavesimulation_amd.torch_backend import adjoint_run

    S = problem.parallel_shots
    engine = problem._engine(S)
    ny, nx = problem.model.shape
    L = int(native.lib().swfd_steps_per_launch(1, ny, nx, 0)) if len(engine.lam) > 3 else 1
    dev = problem.model.device
    R = problem.sampler.n_elements
    grad = torch.zeros(problem.model.shape, device=dev)
    for t in engine.lam[:2]:
        t.zero_()
    kw = dict(state=(engine.lam[0], engine.lam[1]), work=engine.lam,
              image_scale=problem.solver.dt2, grad=grad, workspace=engine.workspace)
    if engine.reconstruct:
        n_steps = 16 * L
        hi, lo = engine.n_steps, engine.n_steps - n_steps
        rows, cols = problem._rows_cols(np.arange(S))
        kw.update(recon=(engine.work[1], engine.work[0]), recon_work=engine.rwork,
                  bnd=engine.ring, sources=(rows, cols, problem.source_values,
                                            problem._source_weights(np.arange(S))))
    else:
        n_steps = max(L, (engine.K // L) * L)  # imaging reads history slots up to n_steps
        hi, lo = n_steps, 0
        kw.update(hist=engine.hist, hist_first=-2, hist_every=1)
    residual = torch.randn((engine.n_steps if engine.reconstruct else n_steps, S, R),
                           device=dev) * 1e-3
    per_run = n_steps // L
    runs = -(-min_launches // per_run)
    adjoint_run(problem.solver, problem.sampler, S, residual, hi, lo, **kw)  # warm
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(runs):
        adjoint_run(problem.solver, problem.sampler, S, residual, hi, lo, **kw)
    stop.record(stream)
    torch.cuda.synchronize()
    n_launch = runs * per_run
    return start.elapsed_time(stop) / 1e3 / n_launch, S, L, n_launch


def adjoint_roofline(problem, t_launch, S, L, n_launch):
    """Roofline entry of the fused adjoint launch (see the module docstring)."""
    import soundwavesimulation_amd._native as native

    ny, nx = problem.model.shape
    R = problem.sampler.n_elements
    cells = ny * nx
    recon = problem._engine(S).reconstruct
    shape = native.adjoint_launch_shape(ny, nx, S, recon)
    if recon:
        # per shot the adjoint pair and the forward pair read and written (8
        # fields), the boundary cells of the L rebuilt fields; c^2 and the
        # gradient (read + write) once; L residual rows
        alg_bytes = 4.0 * (8 * S * cells + 3 * cells + L * S * R +
                           L * S * (2 * nx + 2 * (ny - 2)))
    else:
        # per shot 2 state fields and L forward fields read, 2 fields written;
        # c^2 and the gradient (read + write) once; L residual rows
        alg_bytes = 4.0 * (S * cells * (L + 4) + 3 * cells + L * S * R)
    achieved = alg_bytes / t_launch / 1e9
    if recon:
        kernel = f"adjr_kernel<{L}, every level imaged>"
    elif shape["full_width"]:
        kernel = f"adjm_kernel<{L}, imaging>"
    elif L == 2:
        kernel = "adj2_kernel<residual+imaging>"
    else:
        kernel = "step_kernel<adjoint, residual+imaging>"
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": kernel, "launch_ms": round(t_launch * 1e3, 4),
            "launches_timed": n_launch, "algorithmic_bytes": alg_bytes,
            "launch_shape": {"grid": [ny, nx], "shots": S, **shape}}
    tpath = os.path.join(HERE, "profiles",
                         f"traffic_adjoint{'_recon' if recon else ''}_{ny}x{nx}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        want = {"grid": [ny, nx], "steps_per_launch": L, "shots_per_launch": S,
                "segment_rows": shape["segment_rows"], "shot_groups": shape["shot_groups"]}
        have = {k: tr.get(k) for k in want}
        want_kernel = f"adjr_kernel<{L}," if recon else f"adjm_kernel<{L},"
        if have == want and tr.get("kernel", "").find(want_kernel) >= 0:
            traffic = float(tr["bytes_per_launch"])
            roof["traffic"] = traffic
            roof["traffic_gbs"] = round(traffic / t_launch / 1e9, 1)
            roof["traffic_frac"] = round(traffic / t_launch / 1e9 / HBM_PEAK_GBS, 4)
            roof["traffic_over_algorithmic"] = round(traffic / alg_bytes, 3)
            roof["profile_launch_us"] = tr.get("avg_launch_us_trace")
        else:
            roof["traffic_error"] = (f"{os.path.basename(tpath)} was measured for {have}, "
                                     f"this launch is {want}")
    else:
        roof["traffic_error"] = f"no traffic profile for a {ny}x{nx} grid"
    units = S * cells * L
    roof["algorithmic_equiv"] = {
        "bytes_per_cell_update": 20, "gbs": round(20.0 * units / t_launch / 1e9, 1),
        "frac": round(20.0 * units / t_launch / 1e9 / HBM_PEAK_GBS, 4),
        "note": "single-step streams (read lam, lam_old, c^2, forward field; write lam_next) "
                "times the launch's cell-updates; a multi-step launch moves fewer bytes"}
    return roof


def time_forward_kernel(problem, n_steps=20):
    """Average time per forward step (two steps per launch with four work fields)."""
    from soundwavesimulation_amd.torch_backend import forward_run

    S = problem.parallel_shots
    engine = problem._engine(S)
    rows, cols = problem._rows_cols(np.arange(S))
    kw = dict(work=engine.work, final=False)
    forward_run(problem.solver, S, rows, cols, problem.source_values, n_steps, **kw)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    forward_run(problem.solver, S, rows, cols, problem.source_values, n_steps, **kw)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / 1e3 / n_steps, S


def cpu_baseline(a, c_true, dt, h):
    """Float64 C oracle gradient on the full grid, a few shots and steps."""
    import oracle
    from soundwavesimulation_amd import RingArray, axis_centers, ricker

    oracle.build()
    n = a.grid
    x = axis_centers(n, h)
    array = RingArray(a.shots, a.radius_mm * 1e-3, x, x)
    r0, c0, w = array.footprint()
    shots = np.arange(a.cpu_shots) * (a.shots // max(a.cpu_shots, 1))
    pulse = ricker(dt, a.freq_khz * 1e3)
    N = a.cpu_steps
    obs = np.zeros((len(shots), a.shots, N))
    threads = oracle.n_threads()
    t0 = time.perf_counter()
    oracle.gradient(np.full_like(c_true, 1500.0), h, h, dt, array.rows[shots], array.cols[shots],
                    pulse, N, r0, c0, w, obs, shots)
    sec = time.perf_counter() - t0
    cells = 2.0 * n * n * N * len(shots)
    return {"value": cells / sec / 1e9, "unit": "Gcell/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"float64 C oracle gradient, {n}x{n} grid, {len(shots)} shots, {N} steps "
                      f"(forward+adjoint), {sec:.1f} s"}


def cpu_model():
    """The host CPU's model name, with its family / model / stepping numbers
    (``/proc/cpuinfo``; a virtual machine may report only a generic name)."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if info:
                        break  # the first processor's block is enough
                    continue
                key, value = (t.strip() for t in line.split(":", 1))
                info.setdefault(key, value)
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    ids = [f"{k} {info[k]}" for k in ("cpu family", "model", "stepping") if k in info]
    return f"{name} ({', '.join(ids)})" if ids else name


# The reference's batched op sequence (fdtd2d/torch_backend.py:477-598), restated
# as pure functions of tensors so they can run eagerly or under torch.compile
# (the reference compiles each with mode="reduce-overhead", fullgraph=True,
# :630-638). Comparator only: the product path is the native library.

def _ref_mur(u, un, kl, kr, kb, kt, cr, cc, cri, cci, ckx, cky):
    un[:, 1:-1, 0] = u[:, 1:-1, 1] + kl * (un[:, 1:-1, 1] - u[:, 1:-1, 0])
    un[:, 1:-1, -1] = u[:, 1:-1, -2] + kr * (un[:, 1:-1, -2] - u[:, 1:-1, -1])
    un[:, 0, 1:-1] = u[:, 1, 1:-1] + kb * (un[:, 1, 1:-1] - u[:, 0, 1:-1])
    un[:, -1, 1:-1] = u[:, -2, 1:-1] + kt * (un[:, -2, 1:-1] - u[:, -1, 1:-1])
    cu = u[:, cr, cc]
    fx = u[:, cr, cci] + ckx * (un[:, cr, cci] - cu)
    fy = u[:, cri, cc] + cky * (un[:, cri, cc] - cu)
    un[:, cr, cc] = 0.5 * (fx + fy)
    return un


def _ref_forward_record(u_old, u, c2, kernel, value, bi, rows, cols, boundary, receivers,
                        dt2: float):
    import torch.nn.functional as F

    lap = F.conv2d(u[:, None], kernel, padding=1)[:, 0]
    un = 2.0 * u - u_old + dt2 * c2 * lap
    un[bi, rows, cols] += value
    un = _ref_mur(u, un, *boundary)
    row0, row1, col0, col1, w00, w10, w01, w11 = receivers
    samples = (un[:, row0, col0] * w00 + un[:, row0, col1] * w10 + un[:, row1, col0] * w01
               + un[:, row1, col1] * w11)
    return u, un, samples


def _ref_adjoint_image(u_old, u, c2, kernel, residual, bi, boundary, receivers, forward,
                       gradient, dt2: float, scale: float):
    import torch.nn.functional as F

    row0, row1, col0, col1, w00, w10, w01, w11 = receivers
    src = torch.zeros_like(u)
    shots = bi[:, None]
    for rr, ccc, ww in ((row0, col0, w00), (row0, col1, w10), (row1, col0, w01),
                        (row1, col1, w11)):
        src.index_put_((shots, rr[None], ccc[None]), residual * ww, accumulate=True)
    lap = F.conv2d((c2 * u)[:, None], kernel, padding=1)[:, 0]
    un = 2.0 * u - u_old + dt2 * lap + src
    un = _ref_mur(u, un, *boundary)
    lapf = F.conv2d(forward.to(u.dtype)[:, None], kernel, padding=1)[:, 0]
    return u, un, gradient + scale * (un * lapf).sum(dim=0)


def torch_reference(a, problem, n_fwd=10, device=None, S=8, compiled=False):
    """The reference op sequence per gradient (seconds): S = 8 shots per batch
    (what the reference's memory model admits at 1024^2) on the GPU (default)
    or on the host cores (device="cpu"); ``compiled`` wraps the step functions
    in torch.compile as the reference does on CUDA/ROCm devices (outputs
    cloned after each call, as its batched solver does)."""
    dev = problem.model.device if device is None else torch.device(device)
    sync = torch.cuda.synchronize if dev.type == "cuda" else (lambda: None)
    solver = problem.solver
    sampler = problem.sampler
    ny, nx = problem.model.shape
    c2 = problem.model.c2.to(dev)
    kernel = problem.model.laplacian_kernel.to(dev)
    dt2 = solver.dt2
    boundary = tuple(t.to(dev) for t in (solver.k_left, solver.k_right, solver.k_bottom,
                                         solver.k_top))
    boundary += (torch.tensor([0, 0, -1, -1], device=dev), torch.tensor([0, -1, 0, -1], device=dev),
                 torch.tensor([1, 1, -2, -2], device=dev), torch.tensor([1, -2, 1, -2], device=dev),
                 solver.corner_kx.to(dev), solver.corner_ky.to(dev))
    rows = torch.as_tensor(problem.array.rows[:S], device=dev)
    cols = torch.as_tensor(problem.array.cols[:S], device=dev)
    bi = torch.arange(S, device=dev)
    receivers = tuple(t.to(dev) for t in sampler.receivers)
    fwd_fn, adj_fn = _ref_forward_record, _ref_adjoint_image
    if compiled:
        fwd_fn = torch.compile(fwd_fn, mode="reduce-overhead", fullgraph=True)
        adj_fn = torch.compile(adj_fn, mode="reduce-overhead", fullgraph=True)
    state = [torch.zeros((S, ny, nx), device=dev), torch.zeros((S, ny, nx), device=dev)]
    hist = torch.empty((n_fwd, S, ny, nx), device=dev)
    val = problem.source_values.to(dev)
    g = [torch.zeros((ny, nx), device=dev)]
    res = torch.randn((S, sampler.n_elements), device=dev)

    def fwd(k):
        u_old, u, smp = fwd_fn(state[0], state[1], c2, kernel, val[k], bi, rows, cols,
                               boundary, receivers, dt2)
        if compiled:
            u_old, u, smp = u_old.clone(), u.clone(), smp.clone()
        hist[k].copy_(u)
        state[0], state[1] = u_old, u
        return smp

    def adj(k):
        u_old, u, grad = adj_fn(state[0], state[1], c2, kernel, res, bi, boundary, receivers,
                                hist[k], g[0], dt2, dt2)
        if compiled:
            u_old, u, grad = u_old.clone(), u.clone(), grad.clone()
        state[0], state[1], g[0] = u_old, u, grad

    with torch.inference_mode():
        t_c = time.perf_counter()
        for k in range(min(3, n_fwd)):  # warm-up (compiles and records the graphs)
            fwd(k)
            adj(k)
        sync()
        t_compile = time.perf_counter() - t_c
        t0 = time.perf_counter()
        for k in range(n_fwd):
            fwd(k)
        sync()
        t_f = (time.perf_counter() - t0) / n_fwd
        t0 = time.perf_counter()
        for k in range(n_fwd):
            adj(k)
        sync()
        t_a = (time.perf_counter() - t0) / n_fwd
    per_grad = (t_f + t_a) * a.nsteps * (a.shots / S)
    del hist, state
    if dev.type == "cuda":
        torch.cuda.empty_cache()
    return per_grad, t_f, t_a, S, t_compile


def reference_entry(per_grad, t_f, t_a, S, t_warm, cells, sec, a, label):
    return {"seconds_per_gradient": round(per_grad, 2),
            "gcell_s": round(cells / per_grad / 1e9, 3),
            "speedup": round(per_grad / sec, 2),
            "note": f"{label}, {S} shots per batch, fwd {t_f*1e3:.2f} ms + adj {t_a*1e3:.2f} ms "
                    f"per step, scaled to {a.shots} shots x {a.nsteps} steps; warm-up "
                    f"{t_warm:.1f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              "(shots would shard across GPUs with one all-reduce of the gradient).",
              file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        n = 2 ** 29
        x = torch.ones(n, device="cuda")
        y = torch.ones(n, device="cuda")
        z = torch.empty(n, device="cuda")
        for _ in range(5):
            torch.add(x, y, out=z)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"read_bytes": 2 * 4 * n, "write_bytes": 4 * n,
                                               "launches": 5}}))
        return
    if a.dry_run:
        c_true, x, dt, h, radius, pulse, array = host_geometry(a)
        print(json.dumps({"grid": a.grid, "dt": dt, "radius_m": radius, "shots": a.shots,
                          "pulse_samples": int(pulse.size), "c_range": [float(c_true.min()),
                                                                         float(c_true.max())]}))
        return
    import soundwavesimulation_amd._native as native

    native.load_library()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t_setup = time.perf_counter()
    problem, m0, c_true, dt, h = build_problem(a, dev)
    print(f"# setup {time.perf_counter() - t_setup:.1f} s; batches of {problem.parallel_shots} "
          f"shots; checkpoint every {problem.checkpoint_every} steps", file=sys.stderr,
          flush=True)
    for i in range(a.warmup):
        J, _ = problem.misfit_and_grad(m0)
        print(f"# warmup {i}: J = {J:.6e}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        J, grad = problem.misfit_and_grad(m0)
        print(f"# step {i}: J = {J:.6e}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / a.steps
    n = a.grid
    cells = 2.0 * n * n * a.nsteps * a.shots
    gcells = cells / sec / 1e9

    t_adj, S, L, n_launch = time_adjoint_kernel(problem)
    t_fwd, _ = time_forward_kernel(problem)
    roofline = adjoint_roofline(problem, t_adj, S, L, n_launch)
    roofline["forward_step_ms"] = round(t_fwd * 1e3, 4)
    print(f"# adjoint kernel {t_adj*1e3:.3f} ms/launch of {L} steps over {n_launch} launches "
          f"({roofline['achieved']:.0f} GB/s algorithmic, frac {roofline['frac']:.3f}); "
          f"forward {t_fwd*1e3:.3f} ms/step", file=sys.stderr, flush=True)

    extra = {}
    if not a.no_torch_reference:
        for key, compiled, label in (("torch_reference", False, "eager PyTorch reference op "
                                      "sequence"),
                                     ("torch_reference_compiled", True, "reference op sequence "
                                      "under torch.compile(mode='reduce-overhead', "
                                      "fullgraph=True)")):
            try:
                per_grad, t_f, t_a, S_ref, t_warm = torch_reference(a, problem, compiled=compiled)
                extra[key] = reference_entry(per_grad, t_f, t_a, S_ref, t_warm, cells, sec, a,
                                             label)
            except Exception as exc:  # report, never hide the main number
                extra[key] = {"error": repr(exc)[:300]}
            print(f"# {key}: {extra[key]}", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(a, c_true, dt, h)
        try:
            # the reference's own path on the host cores: the same op sequence
            # on device="cpu", 2 shots x 3 steps each way, scaled to a gradient
            per_grad, t_f, t_a, S_ref, _ = torch_reference(a, problem, n_fwd=3, device="cpu",
                                                           S=2)
            extra["torch_reference_cpu"] = {
                "seconds_per_gradient": round(per_grad, 1),
                "gcell_s": round(cells / per_grad / 1e9, 4),
                "threads": torch.get_num_threads(),
                "sample": f"eager PyTorch reference op sequence on the host, {S_ref} shots, "
                          f"fwd {t_f*1e3:.1f} ms + adj {t_a*1e3:.1f} ms per step, scaled to "
                          f"{a.shots} shots x {a.nsteps} steps"}
        except Exception as exc:  # report, never hide the main number
            extra["torch_reference_cpu"] = {"error": repr(exc)[:200]}
    line = {
        "metric": f"FWI gradient Gcell/s @{a.grid}^2, {a.shots} shots "
                  "(forward+adjoint cell-updates/s)",
        "value": round(gcells, 3), "unit": "Gcell/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(sec * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"fwi_gradient_{n}x{n}_{a.shots}shots_{a.nsteps}steps",
                   "grid": n, "spacing_mm": a.spacing_mm, "shots": a.shots,
                   "time_steps": a.nsteps, "receivers": a.shots, "ricker_khz": a.freq_khz,
                   "checkpoint_every": problem.checkpoint_every,
                   "shots_per_batch": problem.parallel_shots},
        "gradient_seconds": round(sec, 3),
        "gradient_algorithmic_equiv": {
            "bytes_per_cell_update": 20,
            "frac": round(20.0 * cells / sec / 1e9 / HBM_PEAK_GBS, 4),
            "note": "the whole gradient at the single-step 20 B per cell-update model; "
                    "checkpointed multi-step launches move fewer bytes, so this can pass 1"},
        "roofline": roofline, "cpu_baseline": cpu, **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
