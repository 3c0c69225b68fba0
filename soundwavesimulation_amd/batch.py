"""Many independent problems in lockstep: batched launches, HIP graphs, tuning.

The reference batches the shots of one problem (``torch_backend.py:601-770``);
here several problems -- each with its own grid, medium, ring, shot count and
step count -- advance together, and at every tick the multi-step launches of
all of them go out as one batched kernel launch per kernel variant and row
width (``swfd_forward_batch`` / ``swfd_adjoint_batch`` in ``include/swfd.h``):
full-width launches of up to four steps where a row fits, two-step blocks
otherwise.

* ``BatchedForward`` / ``BatchedAdjoint``: one prepared lockstep run over a
  list of per-problem run descriptions (the keyword arguments of
  ``torch_backend.forward_run`` / ``adjoint_run``). ``prepare()`` writes the
  per-tick parameter tables to the device; ``launch()`` only enqueues kernels,
  so it can be captured in a graph.
* ``UnevenGradients``: the checkpointed FWI gradient of every problem (forward
  with checkpoints, residual and misfit, then segment recomputation and
  adjoint with imaging), scheduled so that every problem's segments run
  back to front and the phase of one problem never waits for another's; the
  whole sequence is built once and can be captured into a
  ``torch.cuda.CUDAGraph`` (a HIP graph on ROCm) and replayed.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from .inversion.fwi import (RECON_EVERY, default_checkpoint_every, reconstruction_available,
                            segment_plan)
from .torch_backend import _source_values_tensor, adjoint_run, boundary_floats

_F32 = torch.float32


def _ptr(t):
    return _native.ptr(t)


def _forward_args(solver, n_shots, rows, cols, values, n_steps, *, init, work, samples=None,
                  hist=None, hist_first=0, hist_every=1, ckpt=None, ckpt_every=0, step_begin=0,
                  final=None, bnd=None):
    return _native.ForwardArgs(
        n_shots=n_shots, step_begin=step_begin, n_steps=n_steps,
        init_old=_ptr(init[0]), init_u=_ptr(init[1]),
        work0=_ptr(work[0]), work1=_ptr(work[1]), work2=_ptr(work[2]),
        work3=_ptr(work[3]) if len(work) > 3 else None,
        work4=_ptr(work[4]) if len(work) > 5 else None,
        work5=_ptr(work[5]) if len(work) > 5 else None,
        final_old=_ptr(final[0]) if final is not None else None,
        final_u=_ptr(final[1]) if final is not None else None,
        src_rows=_ptr(rows), src_cols=_ptr(cols), src_values=_ptr(values),
        samples=_ptr(samples), hist=_ptr(hist), hist_first=hist_first, hist_every=hist_every,
        hist_slots=0 if hist is None else hist.shape[0], ckpt=_ptr(ckpt), ckpt_every=ckpt_every,
        ckpt_slots=0 if ckpt is None else ckpt.shape[0],
        bnd=_ptr(bnd), bnd_first=0, bnd_slots=0 if bnd is None else bnd.shape[0])


def _adjoint_args(n_shots, residual, step_hi, step_lo, *, state, work, hist=None, hist_first=0,
                  hist_every=1, image_scale=0.0, grad=None, workspace=None, recon=None,
                  recon_work=None, bnd=None, sources=None):
    rk = {}
    if recon is not None:  # the forward fields rebuilt in the launches (see adjoint_run)
        rows, cols, values = sources
        rk = dict(recon_next=_ptr(recon[0]), recon_cur=_ptr(recon[1]),
                  recon_work0=_ptr(recon_work[0]), recon_work1=_ptr(recon_work[1]),
                  recon_work2=_ptr(recon_work[2]), recon_work3=_ptr(recon_work[3]),
                  bnd=_ptr(bnd), bnd_first=0, bnd_slots=bnd.shape[0], src_rows=_ptr(rows),
                  src_cols=_ptr(cols), src_values=_ptr(values))
    return _native.AdjointArgs(
        n_shots=n_shots, step_hi=step_hi, step_lo=step_lo,
        init_old=_ptr(state[0]), init_u=_ptr(state[1]),
        work0=_ptr(work[0]), work1=_ptr(work[1]), work2=_ptr(work[2]),
        work3=_ptr(work[3]) if len(work) > 3 else None,
        final_old=_ptr(state[0]), final_u=_ptr(state[1]), residual=_ptr(residual),
        hist=_ptr(hist), hist_first=hist_first, hist_every=hist_every,
        hist_slots=0 if hist is None else hist.shape[0], image_scale=float(image_scale),
        grad=_ptr(grad), workspace=_ptr(workspace), **rk)


class _Prepared:
    """Common part of the batched forward / adjoint runs."""

    kind = ""

    def __init__(self, solvers, samplers, args, device):
        n = len(args)
        self.n = n
        self.device = device
        self._keep = (solvers, samplers)
        for s in solvers:
            s._sync_medium()
        self.grids = (_native.Grid * n)(*[s._grid for s in solvers])
        self.media = (_native.Medium * n)(*[s._medium for s in solvers])
        recs = [smp._rec if smp is not None else _native.Receivers() for smp in samplers]
        self.recs = (_native.Receivers * n)(*recs)
        args_t = _native.ForwardArgs if self.kind == "forward" else _native.AdjointArgs
        self.args = (args_t * n)(*args)
        lib = _native.lib()
        nbytes = ctypes.c_size_t(0)
        _native.check(getattr(lib, f"swfd_{self.kind}_batch_table_bytes")(
            n, self.grids, self.media, self.recs, self.args, ctypes.byref(nbytes)))
        self.table_bytes = int(nbytes.value)
        self.table = torch.empty(max(self.table_bytes, 256), dtype=torch.uint8, device=device)

    def prepare(self):
        """Write the per-tick tables to the device (blocking)."""
        _native.check(getattr(_native.lib(), f"swfd_{self.kind}_batch_prepare")(
            self.n, self.grids, self.media, self.recs, self.args, _ptr(self.table),
            self.table.numel()))
        return self

    def launch(self):
        """Enqueue the run on torch's current stream (no host synchronisation)."""
        _native.check(getattr(_native.lib(), f"swfd_{self.kind}_batch")(
            self.n, self.grids, self.media, self.recs, self.args, _ptr(self.table),
            self.table.numel(), _native.stream_of(self.device)))


class BatchedForward(_Prepared):
    """Lockstep forward runs; ``runs`` = list of (solver, sampler or None, kwargs).

    The prepared tables hold raw device pointers, so the object keeps every
    tensor of ``runs`` alive."""

    kind = "forward"

    def __init__(self, runs):
        self._runs = list(runs)
        solvers = [r[0] for r in runs]
        samplers = [r[1] for r in runs]
        args = [_forward_args(r[0], **r[2]) for r in runs]
        super().__init__(solvers, samplers, args, solvers[0].model.device)


class BatchedAdjoint(_Prepared):
    """Lockstep adjoint runs; ``runs`` = list of (solver, sampler, kwargs) (kept
    alive, as for ``BatchedForward``)."""

    kind = "adjoint"

    def __init__(self, runs):
        self._runs = list(runs)
        solvers = [r[0] for r in runs]
        samplers = [r[1] for r in runs]
        args = [_adjoint_args(**r[2]) for r in runs]
        super().__init__(solvers, samplers, args, solvers[0].model.device)


class GradientProblem:
    """One FWI gradient problem of an uneven batch.

    Every ring element fires unless ``sources`` names the transmitting
    elements (sorted, as the reference's shot sampling keeps them,
    ``invert_ring.py:579-590``); ``observed`` holds one gather per element
    ``(n_elements, R, N)``, and each shot's own receiver channel is muted.
    ``reconstruct`` (default: wherever rows fit a full-width launch and
    SWFD_RECON is not 0): the adjoint rebuilds the forward fields backwards
    in time from checkpoints every ``RECON_EVERY`` steps and a ring of the
    boundary cells of every step, as ``CheckpointedGradient(...,
    reconstruct=True)`` does; otherwise segments are recomputed into a field
    history.
    """

    def __init__(self, solver, sampler, array, pulse, n_steps, observed, checkpoint_every=None,
                 sources=None, reconstruct=None):
        self.solver, self.sampler, self.array = solver, sampler, array
        dev = solver.model.device
        self.sources = (np.arange(array.n_elements) if sources is None
                        else np.asarray(sources, dtype=np.int64))
        if self.sources.ndim != 1 or self.sources.size < 1 or \
                self.sources.min() < 0 or self.sources.max() >= array.n_elements:
            raise ValueError("sources must name ring elements.")
        if reconstruct is None:
            reconstruct = reconstruction_available(solver.model.shape)
        elif reconstruct and not reconstruction_available(solver.model.shape):
            raise ValueError("this grid has no reconstructing adjoint (full-width launches).")
        self.reconstruct = bool(reconstruct)
        self.S = int(self.sources.size)
        self.N = int(n_steps)
        default = RECON_EVERY if self.reconstruct else default_checkpoint_every(self.N)
        self.K = int(checkpoint_every or default)
        self.n_seg, n_ckpt, n_hist = segment_plan(self.N, self.K)
        shape = (self.S, *solver.model.shape)
        e = lambda *s: torch.empty(s, device=dev, dtype=_F32)  # noqa: E731
        self.work = [e(*shape) for _ in range(4)]
        self.ckpt = e(max(n_ckpt, 2), *shape) if n_ckpt > 0 else None
        if self.reconstruct:
            # forward: work0..3 and two adjoint fields for the intermediate
            # steps (the final pair stays in work0 / work1); adjoint: four
            # fields for its state, four for the rebuilt forward pairs
            self.lam = [e(*shape) for _ in range(4)]
            self.rwork = [e(*shape) for _ in range(4)]
            self.ring = e(self.N, self.S, boundary_floats(solver.model.shape))
            self.hist = None
        else:
            # The adjoint state shares the forward work fields: the forward is
            # over before the first adjoint step, and the segment recomputation
            # writes every output into the history (its work fields stay idle).
            self.lam = self.work
            self.rwork = []
            self.ring = None
            self.hist = e(n_hist, *shape)
        ny, nx = solver.model.shape
        self.workspace = e(int(_native.lib().swfd_adjoint_workspace_floats(ny, nx, self.S)))
        R = sampler.n_elements
        self.samples = e(self.N, self.S, R)
        self.residual = e(self.N, self.S, R)
        self.res_ws = e(int(_native.lib().swfd_residual_workspace_floats(self.N, self.S, R)) + 1)
        self.misfit = e(1)
        self.grad = torch.zeros(solver.model.shape, device=dev, dtype=_F32)
        rc = [array.inject_rows_cols(int(i)) for i in self.sources]
        self.rows = torch.as_tensor(np.array([r for r, _ in rc], dtype=np.int32), device=dev)
        self.cols = torch.as_tensor(np.array([c for _, c in rc], dtype=np.int32), device=dev)
        self.index = torch.as_tensor(self.sources.astype(np.int32), device=dev)
        self.values = _source_values_tensor(torch.as_tensor(pulse), self.N, dev)
        self.observed = observed.contiguous()
        self.cells = 2.0 * ny * nx * self.N * self.S
        held = [*self.work, self.workspace, self.samples, self.residual]
        held += [self.hist] if self.hist is not None else [*self.lam, *self.rwork, self.ring]
        held += [self.ckpt] if self.ckpt is not None else []
        self.memory_bytes = sum(t.numel() * 4 for t in held)

    def segment(self, j):
        n0 = j * self.K
        return n0, min(n0 + self.K, self.N)

    def residual_and_misfit(self):
        """Residual (N, S, R) and misfit into preallocated buffers (capturable)."""
        lib = _native.lib()
        R = self.sampler.n_elements
        obs = self.observed
        _native.check(lib.swfd_residual(
            self.N, self.S, R, _ptr(self.samples), _ptr(obs), obs.stride(0), obs.stride(1),
            _ptr(self.index), _ptr(self.index), _ptr(self.residual), _ptr(self.res_ws),
            _ptr(self.misfit), _native.stream_of(self.solver.model.device)))


class UnevenGradients:
    """FWI gradients of many uneven problems, batched per tick.

    Schedule: one batched forward over all problems (each for its own step
    count, with samples and checkpoints, and the boundary ring where the
    adjoint rebuilds the forward fields); the residual of each problem; then
    the adjoints:

    * recomputing problems: phase i = 0, 1, ... runs, for every problem with a
      segment left, its segment ``n_seg - 1 - i`` (a checkpoint copy, a
      batched recomputation and a batched adjoint over the recomputed
      fields); short problems finish after fewer phases;
    * reconstructing problems: each problem's segments back to front, the
      problems spread over ``streams`` HIP streams (longest first, each to
      the stream with the least work so far) that run side by side. Their
      launches take their parameters as kernel arguments: the batched
      reconstructing launch, which reads them from a table, holds them in
      registers and spills (433 vector registers to scratch against 66), and
      ran the 24-problem workload at 174 Gcell/s.

    Every problem of a batch reconstructs, or none does.
    """

    def __init__(self, problems, streams=4):
        self.problems = list(problems)
        modes = {p.reconstruct for p in self.problems}
        if len(modes) != 1:
            raise ValueError("the problems of a batch must all reconstruct or all recompute.")
        self.reconstruct = modes.pop()
        self.phases = max(p.n_seg for p in self.problems)
        fwd = []
        for p in self.problems:
            kw = dict(n_shots=p.S, rows=p.rows, cols=p.cols, values=p.values, n_steps=p.N,
                      init=(p.work[0], p.work[1]), samples=p.samples, ckpt=p.ckpt,
                      ckpt_every=p.K if p.ckpt is not None else 0)
            if self.reconstruct:
                # the adjoint's fields keep the intermediate steps of four-step
                # launches; the final pair stays in work0 / work1
                kw.update(work=p.work + [p.lam[0], p.lam[1]], final=(p.work[0], p.work[1]),
                          bnd=p.ring)
            else:
                # the history is idle during the forward pass: two of its fields keep
                # the intermediate steps of four-step launches until their samples are
                # gathered
                kw.update(work=p.work + ([p.hist[0], p.hist[1]] if p.hist.shape[0] >= 2 else []))
            fwd.append((p.solver, p.sampler, kw))
        self.forward = BatchedForward(fwd).prepare()
        self.recompute, self.adjoint, self.members = [], [], []
        if self.reconstruct:
            self._plan_streams(streams)
            return
        for i in range(self.phases):
            mem = [p for p in self.problems if p.n_seg - 1 - i >= 0]
            fwd_runs, adj_runs = [], []
            for p in mem:
                j = p.n_seg - 1 - i
                n0, n1 = p.segment(j)
                n_re = n1 - 1 - n0
                if n_re > 0:
                    fwd_runs.append((p.solver, None, dict(
                        n_shots=p.S, rows=p.rows, cols=p.cols, values=p.values, n_steps=n_re,
                        init=(p.hist[0], p.hist[1]), work=p.work, hist=p.hist,
                        hist_first=n0 - 2, hist_every=1, step_begin=n0)))
                adj_runs.append((p.solver, p.sampler, dict(
                    n_shots=p.S, residual=p.residual, step_hi=n1, step_lo=n0,
                    state=(p.lam[0], p.lam[1]), work=p.lam, hist=p.hist, hist_first=n0 - 2,
                    hist_every=1, image_scale=p.solver.dt2, grad=p.grad,
                    workspace=p.workspace)))
            self.members.append(mem)
            self.recompute.append(BatchedForward(fwd_runs).prepare() if fwd_runs else None)
            self.adjoint.append(BatchedAdjoint(adj_runs).prepare())

    def _plan_streams(self, n_streams):
        dev = self.problems[0].solver.model.device
        n = max(1, min(int(n_streams), len(self.problems)))
        self.streams = [torch.cuda.Stream(dev) for _ in range(n)]
        load = [0.0] * n
        self.lanes = [[] for _ in range(n)]
        for k in sorted(range(len(self.problems)), key=lambda k: -self.problems[k].cells):
            s = load.index(min(load))
            self.lanes[s].append(k)
            load[s] += self.problems[k].cells

    @staticmethod
    def _reconstructing_adjoint(p):
        """Segments back to front; each rebuilds the forward fields from the
        pair after its last two steps: the final state, else the next
        segment's checkpoint."""
        for j in range(p.n_seg - 1, -1, -1):
            n0, n1 = p.segment(j)
            if n1 == p.N:
                pair = (p.work[1], p.work[0])
            else:
                c = n1 // p.K - 1
                pair = (p.ckpt[2 * c + 1], p.ckpt[2 * c])
            adjoint_run(p.solver, p.sampler, p.S, p.residual, n1, n0, state=(p.lam[0], p.lam[1]),
                        work=p.lam, image_scale=p.solver.dt2, grad=p.grad,
                        workspace=p.workspace, recon=pair, recon_work=p.rwork, bnd=p.ring,
                        sources=(p.rows, p.cols, p.values, None))

    def run(self, events=None):
        """Enqueue every gradient. ``events`` (dict problem index -> torch
        event) are recorded right after each problem's last launch."""
        for p in self.problems:
            p.work[0].zero_()
            p.work[1].zero_()
        self.forward.launch()
        for p in self.problems:
            p.residual_and_misfit()
            p.grad.zero_()
            p.lam[0].zero_()
            p.lam[1].zero_()
        if self.reconstruct:
            cur = torch.cuda.current_stream(self.problems[0].solver.model.device)
            for st, lane in zip(self.streams, self.lanes):
                st.wait_stream(cur)
                with torch.cuda.stream(st):
                    for k in lane:
                        self._reconstructing_adjoint(self.problems[k])
                        if events is not None:
                            events[k].record()
            for st in self.streams:
                cur.wait_stream(st)
            return
        for i in range(self.phases):
            if not self.reconstruct:
                for p in self.members[i]:
                    j = p.n_seg - 1 - i
                    if j == 0:
                        p.hist[0].zero_()
                        p.hist[1].zero_()
                    else:
                        p.hist[0].copy_(p.ckpt[2 * (j - 1)])
                        p.hist[1].copy_(p.ckpt[2 * (j - 1) + 1])
            if self.recompute[i] is not None:
                self.recompute[i].launch()
            self.adjoint[i].launch()
            if events is not None:
                for p in self.members[i]:
                    if p.n_seg - 1 - i == 0:
                        events[self.problems.index(p)].record()

    @property
    def cells(self):
        return sum(p.cells for p in self.problems)


def max_wide_groups(ny, nx):
    """Shot groups (one partial image each) a full-width adjoint launch on an
    ny x nx grid may use: 32 on grids of 1024^2 cells and more, up to
    2^25 / (ny nx) (at most 512) on smaller ones -- the library's rule."""
    return max(32, min(512, (32 << 20) // (ny * nx)))


def gradient_problem_bytes(ny, nx, n_shots, n_receivers, n_steps, checkpoint_every=None,
                           reconstruct=None):
    """Device bytes a ``GradientProblem`` allocates (work fields, checkpoints,
    segment history or rebuilt pairs and boundary ring, shot-group images,
    samples, residual)."""
    if reconstruct is None:
        reconstruct = reconstruction_available((ny, nx))
    K = int(checkpoint_every or (RECON_EVERY if reconstruct else
                                 default_checkpoint_every(n_steps)))
    n_seg, n_ckpt, n_hist = segment_plan(n_steps, K)
    plane = ny * nx
    ckpt = max(n_ckpt, 2) if n_ckpt > 0 else 0
    fields = 4 + ckpt + (8 if reconstruct else n_hist)
    ring = n_steps * boundary_floats((ny, nx)) if reconstruct else 0
    images = max(8, min(max_wide_groups(ny, nx), n_shots))
    return 4 * (fields * n_shots * plane + n_shots * ring + images * plane +
                2 * n_steps * n_shots * n_receivers)


def plan_waves(sizes, budget_bytes, wave_cells):
    """Group problems into waves that run one after the other.

    ``sizes[i] = (work, bytes, cells_per_step)`` of problem i. Problems are taken
    in order of increasing work (shortest first, which minimises the mean
    completion time of a sequence) and a wave is closed when the next problem
    would take its device memory past ``budget_bytes`` or its cells per step
    (the parallel width of its launches) past ``wave_cells``. Returns lists of
    problem indices. Raises if a single problem exceeds the budget.
    """
    order = sorted(range(len(sizes)), key=lambda i: (sizes[i][0], i))
    waves, cur, mem, width = [], [], 0, 0
    for i in order:
        _, nbytes, cells = sizes[i]
        if nbytes > budget_bytes:
            raise ValueError(f"problem {i} needs {nbytes / 1e9:.1f} GB, more than the budget "
                             f"of {budget_bytes / 1e9:.1f} GB")
        if cur and (mem + nbytes > budget_bytes or width + cells > wave_cells):
            waves.append(cur)
            cur, mem, width = [], 0, 0
        cur.append(i)
        mem += nbytes
        width += cells
    if cur:
        waves.append(cur)
    return waves


class WaveSchedule:
    """Uneven gradients in memory-bounded waves, shortest work first.

    ``make_problem(i)`` builds problem i (a ``GradientProblem``; its device
    buffers are allocated then and released after its wave). Each wave is one
    ``UnevenGradients`` batch; waves run back to back on the current stream.
    ``run`` returns the gradients and misfits (host copies) of every problem.
    With ``events`` it records an event after each problem's last launch, and
    ``spans`` (a list) receives a (start, end) event pair around each wave's
    launches. ``graph=True`` runs each wave once eagerly on a side stream,
    captures it as a HIP graph and replays it (the span covers the replay
    only; results equal the eager run). ``graph="whole"`` allocates every
    wave's problems at once (their bytes must fit the budget together),
    captures all the waves in order into one HIP graph and replays that
    graph once: one launch from the host for the whole schedule (``spans``
    gets one pair around the replay; no per-problem events).
    """

    def __init__(self, make_problem, sizes, budget_bytes, wave_cells):
        self.make_problem = make_problem
        self.sizes = list(sizes)
        self.budget = float(budget_bytes)
        self.waves = plan_waves(self.sizes, budget_bytes, wave_cells)

    def total_bytes(self):
        return sum(self.sizes[i][1] for wave in self.waves for i in wave)

    def _run_whole(self, spans=None):
        if self.total_bytes() > self.budget:
            raise ValueError(f"the whole schedule needs {self.total_bytes() / 1e9:.1f} GB, "
                             f"over the {self.budget / 1e9:.1f} GB budget: run it in waves")
        batches = []
        for wave in self.waves:
            problems = [self.make_problem(i) for i in wave]
            batches.append((wave, problems, UnevenGradients(problems)))
        dev = batches[0][1][0].solver.model.device
        cur = torch.cuda.current_stream(dev)
        side = torch.cuda.Stream(dev)
        side.wait_stream(cur)
        with torch.cuda.stream(side):  # warm-up on the capture stream
            for _, _, ug in batches:
                ug.run()
        cur.wait_stream(side)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _, _, ug in batches:
                ug.run()
        span = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        span[0].record()
        g.replay()
        span[1].record()
        if spans is not None:
            spans.append(span)
        results = {}
        for wave, problems, _ in batches:
            for k, i in enumerate(wave):
                results[i] = (problems[k].grad.cpu(), float(problems[k].misfit.cpu()))
        del g, batches
        return results

    def run(self, events=None, graph=False, spans=None):
        if graph == "whole":
            return self._run_whole(spans)
        results = {}
        for wave in self.waves:
            problems = [self.make_problem(i) for i in wave]
            ug = UnevenGradients(problems)
            dev = problems[0].solver.model.device
            span = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            if graph:
                side = torch.cuda.Stream(dev)
                side.wait_stream(torch.cuda.current_stream(dev))
                with torch.cuda.stream(side):  # warm-up on the capture stream
                    ug.run()
                torch.cuda.current_stream(dev).wait_stream(side)
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    ug.run()
                span[0].record()
                g.replay()
                span[1].record()
                if events is not None:  # the graph replays as a whole
                    for i in wave:
                        events[i].record()
            else:
                span[0].record()
                ug.run(None if events is None else {k: events[i] for k, i in enumerate(wave)})
                span[1].record()
            if spans is not None:
                spans.append(span)
            for k, i in enumerate(wave):
                p = problems[k]
                results[i] = (p.grad.cpu(), float(p.misfit.cpu()))
            del ug, problems
        return results


__all__ = ["BatchedForward", "BatchedAdjoint", "GradientProblem", "UnevenGradients",
           "WaveSchedule", "gradient_problem_bytes", "plan_waves"]
