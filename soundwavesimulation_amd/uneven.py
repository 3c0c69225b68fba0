"""Uneven batch of FWI gradient problems on one GPU.

``python -m soundwavesimulation_amd.uneven`` draws 24 independent problems --
grid from {256^2, 384^2, 512^2, 768^2} at 0.5 mm, a ring of 16-128 elements
(every element fires), 1500-5000 steps, a seeded ellipse phantom each (c
1400-1650 m/s) and a 300 kHz Ricker -- simulates their observed data, tunes
the launch shapes of every grid size on the device (``autotune.tune``; the
forward only when the adjoint rebuilds the forward fields), and computes all
24 gradients (from a homogeneous 1500 m/s start; the adjoint rebuilds the
forward fields, or with ``--recompute`` recomputes checkpointed segments)
with ``batch.WaveSchedule``: problems in order of increasing work, grouped
into waves whose device memory stays within the budget (``--budget-gb``, so
any draw runs) and whose cells per step stay within ``--wave-cells``; each
wave is one ``UnevenGradients`` batch (one batched launch per kernel variant
and tick), waves back to back. When every wave's buffers fit the budget
together, the whole schedule is also captured into one HIP graph and timed
as one replay.

Reported (one JSON line): throughput in cell-updates per second (2 ny nx N S
per problem, summed) over the eager run and over per-wave graph replays, the
makespan, and per-problem latency percentiles from the eager run (an event
after each problem's last launch, measured from the start of the first wave;
the set-up of a wave -- buffer allocation and table upload -- lies between
waves and is included).
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from .array import RingArray
from .autotune import tune
from .batch import GradientProblem, WaveSchedule, gradient_problem_bytes
from .experiment import axis_centers, stable_dt
from .inversion.fwi import reconstruction_available
from .phantoms import ellipse_phantom
from .sources import ricker
from .torch_backend import TorchFDTD2D, TorchRingSampler, TorchScalarWave2D, simulate_shots_batch

SIZES = (256, 384, 512, 768)


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def work_of(spec):
    h, n_el, n_steps = spec
    return 2.0 * h * h * n_steps * n_el


def draw(seed, n=24):
    rng = np.random.default_rng(seed)
    H = rng.choice(SIZES, n)
    R = rng.integers(16, 129, n)
    N = rng.integers(1500, 5001, n)
    return [(int(h), int(r), int(s)) for h, r, s in zip(H, R, N)]


def observed_data(spec, index, seed, device, spacing=0.5e-3, f_peak=300e3):
    """Ring, pulse and observed gathers (n_el, n_el, N) of problem ``index``."""
    h, n_el, n_steps = spec
    x = axis_centers(h, spacing)
    radius = 0.43 * h * spacing
    dt = stable_dt(1650.0, spacing, spacing, cfl=0.3)
    pulse = torch.as_tensor(ricker(dt, f_peak), dtype=torch.float32)
    array = RingArray(n_el, radius, x, x)
    sampler = TorchRingSampler(array, device=device)
    c_true = ellipse_phantom(h, h, spacing, 8, 1000 * seed + index, radius)
    truth = TorchFDTD2D(TorchScalarWave2D(c_true, spacing, device=device), dt)
    observed = torch.empty((n_el, n_el, n_steps), device=device)
    per = 64
    for s0 in range(0, n_el, per):
        sl = slice(s0, min(n_el, s0 + per))
        tr, _ = simulate_shots_batch(truth, sampler, pulse.to(device), array.rows[sl],
                                     array.cols[sl], n_steps)
        observed[sl] = tr
    return dict(array=array, sampler=sampler, pulse=pulse, dt=dt, observed=observed,
                spacing=spacing)


def problem_from(spec, data, device, reconstruct=None):
    """The gradient problem of ``spec`` from a homogeneous 1500 m/s start."""
    h, _, n_steps = spec
    solver = TorchFDTD2D(TorchScalarWave2D(np.full((h, h), 1500.0), data["spacing"],
                                           device=device), data["dt"])
    return GradientProblem(solver, data["sampler"], data["array"], data["pulse"], n_steps,
                           data["observed"], reconstruct=reconstruct)


def sizes_of(specs, reconstruct=None):
    return [(work_of(sp), gradient_problem_bytes(sp[0], sp[0], sp[1], sp[1], sp[2],
                                                 reconstruct=reconstruct),
             sp[1] * sp[0] * sp[0]) for sp in specs]


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--problems", type=int, default=24)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--budget-gb", type=float, default=200.0,
                   help="device memory a wave may hold (observed data aside)")
    p.add_argument("--wave-cells", type=float, default=96e6,
                   help="cells per step (shots x grid cells) a wave may hold")
    p.add_argument("--scale", type=float, default=1.0,
                   help="multiply step counts (a quick reduced run: e.g. 0.05)")
    p.add_argument("--no-tune", action="store_true")
    p.add_argument("--recompute", action="store_true",
                   help="recompute checkpointed segments instead of rebuilding the forward "
                        "fields in the adjoint")
    p.add_argument("--no-graph", action="store_true")
    p.add_argument("--out", default=None)
    a = p.parse_args(argv)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    specs = draw(a.seed, a.problems)
    if a.scale != 1.0:
        specs = [(h, r, max(8, int(round(s * a.scale)))) for h, r, s in specs]
    recon = not a.recompute and all(reconstruction_available((h, h)) for h, _, _ in specs)
    sizes = sizes_of(specs, recon)
    log(f"seed {a.seed}: {len(specs)} problems, {sum(z[1] for z in sizes) / 1e9:.0f} GB of "
        "gradient buffers in all")
    tuned = {}
    if not a.no_tune:
        for h in sorted({s[0] for s in specs}):
            shots = max(r for hh, r, _ in specs if hh == h)
            tuned[h] = tune(h, h, shots, dev, use_cache=False, log=log, adjoint=not recon)
    t0 = time.perf_counter()
    data = [observed_data(s, i, a.seed, dev) for i, s in enumerate(specs)]
    torch.cuda.synchronize()
    log(f"observed data in {time.perf_counter() - t0:.1f} s")
    cells = sum(z[0] for z in sizes)
    sched = WaveSchedule(lambda i: problem_from(specs[i], data[i], dev, recon), sizes,
                         a.budget_gb * 1e9, a.wave_cells)
    waves = [[int(i) for i in w] for w in sched.waves]
    log(f"{len(waves)} waves: " + ", ".join(str(len(w)) for w in waves))

    def timed(graph):
        start = torch.cuda.Event(enable_timing=True)
        events = {i: torch.cuda.Event(enable_timing=True) for i in range(len(specs))}
        spans = []
        torch.cuda.synchronize()
        t_wall = time.perf_counter()
        start.record()
        sched.run(events=events, graph=graph, spans=spans)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_wall
        lat = np.array([start.elapsed_time(events[i]) / 1e3 for i in range(len(specs))])
        busy = [a.elapsed_time(b) / 1e3 for a, b in spans]
        # the same on the GPU's launch clock: waves back to back, the set-up
        # between them (allocation, table upload) left out
        offsets = np.cumsum([0.0] + busy[:-1])
        lat_gpu = np.empty(len(specs))
        for w, wave in enumerate(waves):
            w0 = start.elapsed_time(spans[w][0]) / 1e3
            for i in wave:
                lat_gpu[i] = offsets[w] + lat[i] - w0
        return wall, lat, busy, lat_gpu

    timed(False)  # warm-up (kernel loading, allocator)
    wall, lat, busy, lat_gpu = timed(False)
    makespan = float(lat.max())
    log(f"eager: {wall:.2f} s wall, {sum(busy):.2f} s of wave launches, "
        f"{cells / wall / 1e9:.1f} Gcell/s; p50 latency {np.percentile(lat, 50):.2f} s of "
        f"{makespan:.2f} s")
    report = {
        "workload": f"uneven_{len(specs)}_fwi_gradients",
        "seed": a.seed,
        "adjoint": "rebuilds the forward fields" if recon else "recomputed segments",
        "problems": [{"grid": h, "shots": r, "steps": s} for h, r, s in specs],
        "waves": waves,
        "budget_gb": a.budget_gb, "wave_cells": a.wave_cells,
        "cell_updates": cells,
        "eager_seconds": round(wall, 3),
        "gcell_s_eager": round(cells / wall / 1e9, 2),
        "wave_launch_seconds": round(sum(busy), 3),
        "gcell_s_launches": round(cells / sum(busy) / 1e9, 2),
        "makespan_s": round(makespan, 3),
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 3),
                      "p99": round(float(np.percentile(lat, 99)), 3),
                      "min": round(float(lat.min()), 3), "max": round(float(lat.max()), 3)},
        "p50_over_makespan": round(float(np.percentile(lat, 50)) / makespan, 3),
        "latency_launch_clock_s": {"p50": round(float(np.percentile(lat_gpu, 50)), 3),
                                   "p99": round(float(np.percentile(lat_gpu, 99)), 3),
                                   "max": round(float(lat_gpu.max()), 3)},
        "tuned": {str(k): v for k, v in tuned.items()},
    }
    if not a.no_graph:
        _, _, busy_g, _ = timed(True)
        report["graph_replay_seconds"] = round(sum(busy_g), 3)
        report["gcell_s_graph"] = round(cells / sum(busy_g) / 1e9, 2)
        log(f"graphs: {sum(busy_g):.2f} s of replays, {cells / sum(busy_g) / 1e9:.1f} Gcell/s")
        if sched.total_bytes() <= sched.budget:
            # every wave captured into one graph, replayed once
            spans = []
            torch.cuda.synchronize()
            sched.run(graph="whole", spans=spans)
            torch.cuda.synchronize()
            t_whole = spans[0][0].elapsed_time(spans[0][1]) / 1e3
            report["whole_graph_seconds"] = round(t_whole, 3)
            report["gcell_s_whole_graph"] = round(cells / t_whole / 1e9, 2)
            log(f"one graph of all {len(waves)} waves: {t_whole:.2f} s, "
                f"{cells / t_whole / 1e9:.1f} Gcell/s")
    line = json.dumps(report)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
