"""Benchmark: one full-waveform-inversion gradient on one MI355X.

Workload (default): 1024 x 1024 grid at 0.25 mm, 256-element ring of radius
110 mm, all 256 shots per gradient, 6000 steps, 300 kHz Ricker, true medium
of 12 seeded ellipses (c 1400-1650 m/s), start = truth blurred by 10 mm.
A "step" is one call of ``TorchLeastSquaresFWI.misfit_and_grad`` (forward
with checkpoints, residual and misfit, segment recomputation, adjoint with
imaging, host chain rule to slowness squared). Inputs are synthetic and
resident in HBM before the timed region.

Printed JSON (one line): Gcell/s = 2 * ny * nx * steps * shots / seconds per
gradient, plus
  roofline      -- the adjoint kernel (the steps per launch the library
                   reports, four at this size): 20 B per cell-update (read
                   lam, lam_old, c^2, forward field; write lam_next) times
                   the cell-updates of one launch, over its average launch
                   time, timed with events on torch's stream;
  cpu_baseline  -- the float64 C oracle (OpenMP, all host cores) on a
                   bounded sample of the same workload;
  torch_reference -- the reference's op sequence (conv2d Laplacian,
                   elementwise leapfrog, index_put_ sources, Mur slices,
                   bilinear gather, imaging) in eager PyTorch on the same
                   GPU, timed over a few steps and scaled to one gradient.
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


def time_adjoint_kernel(problem, n_steps=20):
    """Average launch time of the fused adjoint kernel (residual + imaging).

    The run advances ``L`` adjoint steps per launch (``swfd_steps_per_launch``),
    so ``n_steps`` (a multiple of L) steps are ``n_steps / L`` launches; the
    fill and reduction of the partial images around them are included (two
    small kernels, well under 1 % here).
    """
    import soundwavesimulation_amd._native as native
    from soundwavesimulation_amd.torch_backend import adjoint_run

    S = problem.parallel_shots
    engine = problem._engine(S)
    ny, nx = problem.model.shape
    L = int(native.lib().swfd_steps_per_launch(1, ny, nx, 0)) if len(engine.lam) > 3 else 1
    n_steps = min(n_steps, engine.K) // L * L  # imaging reads history slots up to n_steps
    n_launch = n_steps // L
    dev = problem.model.device
    R = problem.sampler.n_elements
    residual = torch.randn((n_steps, S, R), device=dev) * 1e-3
    grad = torch.zeros(problem.model.shape, device=dev)
    for t in engine.lam[:2]:
        t.zero_()
    kw = dict(state=(engine.lam[0], engine.lam[1]), work=engine.lam, hist=engine.hist,
              hist_first=-2, hist_every=1, image_scale=problem.solver.dt2, grad=grad)
    adjoint_run(problem.solver, problem.sampler, S, residual, n_steps, 0, **kw)  # warm
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    adjoint_run(problem.solver, problem.sampler, S, residual, n_steps, 0, **kw)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / 1e3 / n_launch, S, n_steps // n_launch


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
            "sample": f"float64 C oracle gradient, {n}x{n} grid, {len(shots)} shots, {N} steps "
                      f"(forward+adjoint), {sec:.1f} s"}


def torch_reference(a, problem, n_fwd=10, device=None, S=8):
    """Reference op sequence in eager PyTorch, per gradient (seconds), on the
    GPU (default; S = 8 shots per batch, what the reference's memory model
    admits at 1024^2) or on the host cores (device="cpu")."""
    import torch.nn.functional as F

    dev = problem.model.device if device is None else torch.device(device)
    sync = torch.cuda.synchronize if dev.type == "cuda" else (lambda: None)
    solver = problem.solver
    sampler = problem.sampler
    ny, nx = problem.model.shape
    c2 = problem.model.c2.to(dev)
    kernel = problem.model.laplacian_kernel.to(dev)
    dt2 = solver.dt2
    kl, kr, kb, kt = (t.to(dev) for t in (solver.k_left, solver.k_right, solver.k_bottom,
                                          solver.k_top))
    ckx, cky = solver.corner_kx, solver.corner_ky
    if torch.is_tensor(ckx):
        ckx, cky = ckx.to(dev), cky.to(dev)
    cr = torch.tensor([0, 0, -1, -1], device=dev)
    cc = torch.tensor([0, -1, 0, -1], device=dev)
    cri = torch.tensor([1, 1, -2, -2], device=dev)
    cci = torch.tensor([1, -2, 1, -2], device=dev)
    rows = torch.as_tensor(problem.array.rows[:S], device=dev)
    cols = torch.as_tensor(problem.array.cols[:S], device=dev)
    bi = torch.arange(S, device=dev)
    row0, row1, col0, col1, w00, w10, w01, w11 = (t.to(dev) for t in sampler.receivers)

    def mur(u, un):
        un[:, 1:-1, 0] = u[:, 1:-1, 1] + kl * (un[:, 1:-1, 1] - u[:, 1:-1, 0])
        un[:, 1:-1, -1] = u[:, 1:-1, -2] + kr * (un[:, 1:-1, -2] - u[:, 1:-1, -1])
        un[:, 0, 1:-1] = u[:, 1, 1:-1] + kb * (un[:, 1, 1:-1] - u[:, 0, 1:-1])
        un[:, -1, 1:-1] = u[:, -2, 1:-1] + kt * (un[:, -2, 1:-1] - u[:, -1, 1:-1])
        cu = u[:, cr, cc]
        fx = u[:, cr, cci] + ckx * (un[:, cr, cci] - cu)
        fy = u[:, cri, cc] + cky * (un[:, cri, cc] - cu)
        un[:, cr, cc] = 0.5 * (fx + fy)
        return un

    def gather(f):
        return (f[:, row0, col0] * w00 + f[:, row0, col1] * w10 + f[:, row1, col0] * w01
                + f[:, row1, col1] * w11)

    u_old = torch.zeros((S, ny, nx), device=dev)
    u = torch.zeros_like(u_old)
    hist = torch.empty((n_fwd, S, ny, nx), device=dev)
    val = problem.source_values.to(dev)
    g = torch.zeros((ny, nx), device=dev)
    res = torch.randn((S, sampler.n_elements), device=dev)

    def fwd(k):
        nonlocal u_old, u
        lap = F.conv2d(u[:, None], kernel, padding=1)[:, 0]
        un = 2.0 * u - u_old + dt2 * c2 * lap
        un[bi, rows, cols] += val[k]
        un = mur(u, un)
        smp = gather(un)
        hist[k].copy_(un)
        u_old, u = u, un
        return smp

    def adj(k):
        nonlocal u_old, u, g
        src = torch.zeros_like(u)
        shots = bi[:, None]
        for rr, ccc, ww in ((row0, col0, w00), (row0, col1, w10), (row1, col0, w01),
                            (row1, col1, w11)):
            src.index_put_((shots, rr[None], ccc[None]), res * ww, accumulate=True)
        lap = F.conv2d((c2 * u)[:, None], kernel, padding=1)[:, 0]
        un = 2.0 * u - u_old + dt2 * lap + src
        un = mur(u, un)
        lapf = F.conv2d(hist[k].to(torch.float32)[:, None], kernel, padding=1)[:, 0]
        g = g + dt2 * (un * lapf).sum(dim=0)
        u_old, u = u, un

    with torch.inference_mode():
        for k in range(min(3, n_fwd)):
            fwd(k)
            adj(k)
        sync()
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
    del hist, u, u_old
    if dev.type == "cuda":
        torch.cuda.empty_cache()
    return per_grad, t_f, t_a, S


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

    t_adj, S, steps_per_launch = time_adjoint_kernel(problem)
    t_fwd, _ = time_forward_kernel(problem)
    units = S * n * n
    # 20 B per adjoint cell-update (read lam, lam_old, c^2, forward field;
    # write lam_next), steps_per_launch steps per launch
    adj_bytes = 20.0 * units * steps_per_launch
    achieved = adj_bytes / t_adj / 1e9
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes of this kernel (same grid and steps per launch), per shot
    traffic = None
    tpath = os.path.join(HERE, "profiles", f"traffic_adjoint_{n}x{n}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        if tr.get("steps_per_launch") == steps_per_launch:
            traffic = tr["bytes_per_launch_per_shot"] * S
    if steps_per_launch > 2 or (steps_per_launch == 2 and native.lib().swfd_lap_history(n, n)):
        kernel = (f"adjm_kernel<{steps_per_launch}, imaging> (full-width, "
                  f"{steps_per_launch} adjoint steps per launch)")
    elif steps_per_launch == 2:
        kernel = "adj2_kernel<residual+imaging> (two adjoint steps per launch)"
    else:
        kernel = "step_kernel<adjoint, residual+imaging>"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kernel, "launch_ms": round(t_adj * 1e3, 4),
                "steps_per_launch": steps_per_launch, "shots_per_launch": S,
                "forward_step_ms": round(t_fwd * 1e3, 4),
                "forward_gbs_16B": round(16.0 * units / t_fwd / 1e9, 1)}
    if traffic is not None:
        # measured HBM bytes over the same launch time: the rate the memory
        # system actually carries (multi-step launches move fewer bytes than
        # the 20 B per cell-update model, so `achieved` can exceed the peak)
        roofline["traffic_gbs"] = round(traffic / t_adj / 1e9, 1)
        roofline["traffic_frac"] = round(traffic / t_adj / 1e9 / HBM_PEAK_GBS, 4)
    print(f"# adjoint kernel {t_adj*1e3:.3f} ms/launch of {steps_per_launch} steps "
          f"({achieved:.0f} GB/s at 20 B/cell-update); forward {t_fwd*1e3:.3f} ms/step",
          file=sys.stderr, flush=True)

    extra = {}
    if not a.no_torch_reference:
        try:
            per_grad, t_f, t_a, S_ref = torch_reference(a, problem)
            extra["torch_reference"] = {
                "seconds_per_gradient": round(per_grad, 2),
                "gcell_s": round(cells / per_grad / 1e9, 3),
                "speedup": round(per_grad / sec, 2),
                "note": f"eager PyTorch reference op sequence, {S_ref} shots per batch, "
                        f"fwd {t_f*1e3:.2f} ms + adj {t_a*1e3:.2f} ms per step, scaled to "
                        f"{a.shots} shots x {a.nsteps} steps"}
        except Exception as exc:  # report, never hide the main number
            extra["torch_reference"] = {"error": repr(exc)[:200]}
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(a, c_true, dt, h)
        try:
            # the reference's own path on the host cores: the same op sequence
            # on device="cpu", 2 shots x 3 steps each way, scaled to a gradient
            per_grad, t_f, t_a, S_ref = torch_reference(a, problem, n_fwd=3, device="cpu", S=2)
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
        "hbm_frac_gradient": round(20.0 * cells / sec / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roofline, "cpu_baseline": cpu, **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
