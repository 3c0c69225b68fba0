"""Batched lockstep runs of uneven problems against separate runs.

Problems of different grids, shot counts, step counts (odd and even) and
checkpoint/history settings advance together in batched launches; samples,
final states, checkpoints, adjoint states and gradients must equal those of
separate ``forward_run`` / ``adjoint_run`` calls bit for bit (same kernels,
same parameters per problem). The uneven gradient driver is checked against
the single-problem checkpointed gradient, and its captured graph against its
eager run over several replays.
"""

import numpy as np
import pytest
import torch

from soundwavesimulation_amd import RingArray, axis_centers, ricker, stable_dt
from soundwavesimulation_amd.batch import (BatchedAdjoint, BatchedForward, GradientProblem,
                                           UnevenGradients)
from soundwavesimulation_amd.inversion.fwi import CheckpointedGradient, residual_and_misfit
from soundwavesimulation_amd.torch_backend import (TorchFDTD2D, TorchRingSampler,
                                                   TorchScalarWave2D, _source_values_tensor,
                                                   adjoint_run, adjoint_workspace,
                                                   boundary_floats, forward_run,
                                                   simulate_shots_batch)

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
F32 = torch.float32
SPECS = [(96, 70, 5, 41), (64, 64, 3, 30), (150, 131, 8, 25)]  # ny, nx, ring elements, steps
# gradient problems: long enough for waves to reach the other receivers
GRAD_SPECS = [(96, 70, 12, 301), (64, 64, 9, 240), (150, 131, 16, 177)]


def _problem(ny, nx, n_el, seed):
    sp = 1e-3
    rng = np.random.default_rng(seed)
    c = 1450.0 + 150.0 * rng.random((ny, nx))
    dt = stable_dt(1650.0, sp, sp, cfl=0.3)
    solver = TorchFDTD2D(TorchScalarWave2D(c, sp, device=DEV), dt)
    x, y = axis_centers(nx, sp), axis_centers(ny, sp)
    array = RingArray(n_el, 0.4 * min(nx, ny) * sp, x, y)
    sampler = TorchRingSampler(array, device=DEV)
    rc = [array.inject_rows_cols(i) for i in range(n_el)]
    rows = torch.tensor([r for r, _ in rc], dtype=torch.int32, device=DEV)
    cols = torch.tensor([c for _, c in rc], dtype=torch.int32, device=DEV)
    return solver, sampler, array, rows, cols, dt


@pytest.mark.parametrize("extra", [0, 2])  # 2: four-step launches + batched level gathers
def test_batched_forward_equals_separate_runs(extra):
    print(f"batched forward, 3 problems, {4 + extra} work fields", flush=True)
    runs, refs = [], []
    for i, (ny, nx, n_el, N) in enumerate(SPECS):
        solver, sampler, array, rows, cols, dt = _problem(ny, nx, n_el, i)
        values = torch.randn(N + 2, generator=torch.Generator().manual_seed(i)).to(DEV)
        shape = (n_el, ny, nx)
        out = {}
        for mode in ("single", "batch"):
            n_work = 4 + (extra if mode == "batch" else 0)
            work = [torch.zeros(shape, device=DEV) for _ in range(n_work)]
            samples = torch.full((N, n_el, n_el), float("nan"), device=DEV)
            ckpt = torch.zeros((6, *shape), device=DEV)
            kw = dict(n_shots=n_el, rows=rows, cols=cols, values=values, n_steps=N,
                      init=(work[0], work[1]), work=work, samples=samples, ckpt=ckpt,
                      ckpt_every=7)
            out[mode] = (work, samples, ckpt, kw)
        w, smp, ck, kw = out["single"]
        forward_run(solver, n_el, rows, cols, values, N, samples=smp, sampler=sampler, work=w,
                    ckpt=ck, ckpt_every=7, final=True)
        refs.append((w, smp, ck))
        w2, smp2, ck2, kw2 = out["batch"]
        kw2["final"] = (w2[0], w2[1])
        runs.append(((solver, sampler, kw2), (w2, smp2, ck2)))
    batched = BatchedForward([r[0] for r in runs]).prepare()
    batched.launch()
    torch.cuda.synchronize()
    for (_, (w2, smp2, ck2)), (w, smp, ck) in zip(runs, refs):
        assert torch.equal(smp2, smp)
        assert torch.equal(ck2, ck)
        assert torch.equal(w2[0], w[0]) and torch.equal(w2[1], w[1])


def test_batched_forward_history_equals_separate_runs():
    """Without samples the batched runs take four steps per launch (full width,
    one batched launch per row width) and keep every step in the history."""
    print("batched forward with history, 3 problems", flush=True)
    runs, refs = [], []
    for i, (ny, nx, n_el, N) in enumerate(SPECS):
        solver, sampler, array, rows, cols, dt = _problem(ny, nx, n_el, 40 + i)
        values = torch.randn(N + 2, generator=torch.Generator().manual_seed(i)).to(DEV)
        shape = (n_el, ny, nx)
        res = {}
        for mode in ("single", "batch"):
            work = [torch.zeros(shape, device=DEV) for _ in range(4)]
            hist = torch.full((N, *shape), float("nan"), device=DEV)
            res[mode] = (work, hist)
        w, hist = res["single"]
        forward_run(solver, n_el, rows, cols, values, N, hist=hist, work=w, final=True)
        refs.append((w, hist))
        w2, hist2 = res["batch"]
        runs.append(((solver, None, dict(n_shots=n_el, rows=rows, cols=cols, values=values,
                                         n_steps=N, init=(w2[0], w2[1]), work=w2, hist=hist2,
                                         final=(w2[0], w2[1]))), (w2, hist2)))
    batched = BatchedForward([r[0] for r in runs]).prepare()
    batched.launch()
    torch.cuda.synchronize()
    for (_, (w2, hist2)), (w, hist) in zip(runs, refs):
        assert bool(torch.isfinite(hist2).all())
        assert torch.equal(hist2, hist)
        assert torch.equal(w2[0], w[0]) and torch.equal(w2[1], w[1])


@pytest.fixture(params=["wide", "blocks"])
def adjoint_form(request):
    """Adjoint form of both the separate and the batched runs: full width (the
    default) or two-step blocks (forced through the tuning table)."""
    from soundwavesimulation_amd import _native

    lib = _native.lib()
    lib.swfd_tune_clear()
    if request.param == "blocks":
        for ny, nx, _, _ in SPECS:
            _native.check(lib.swfd_tune(3, ny, nx, 0, 0, 2))
    yield request.param
    lib.swfd_tune_clear()


def test_batched_adjoint_equals_separate_runs(adjoint_form):
    print(f"batched adjoint ({adjoint_form}), 3 problems", flush=True)
    runs, refs = [], []
    for i, (ny, nx, n_el, N) in enumerate(SPECS):
        solver, sampler, array, rows, cols, dt = _problem(ny, nx, n_el, 10 + i)
        g = torch.Generator().manual_seed(20 + i)
        shape = (n_el, ny, nx)
        residual = torch.randn((N, n_el, n_el), generator=g).to(DEV)
        hist = torch.randn((N + 1, *shape), generator=g).to(DEV)
        init = [torch.randn(shape, generator=g).to(DEV) for _ in range(2)]
        res = {}
        for mode in ("single", "batch"):
            lam = [init[0].clone(), init[1].clone(), torch.empty(shape, device=DEV),
                   torch.empty(shape, device=DEV)]
            grad = torch.zeros((ny, nx), device=DEV)
            ws = adjoint_workspace(solver, n_el)
            res[mode] = (lam, grad, ws)
        lam, grad, ws = res["single"]
        adjoint_run(solver, sampler, n_el, residual, N, 1, state=(lam[0], lam[1]), work=lam,
                    hist=hist, hist_first=-1, hist_every=1, image_scale=1e-3, grad=grad,
                    workspace=ws)
        refs.append((lam, grad))
        lam2, grad2, ws2 = res["batch"]
        runs.append(((solver, sampler, dict(n_shots=n_el, residual=residual, step_hi=N,
                                            step_lo=1, state=(lam2[0], lam2[1]), work=lam2,
                                            hist=hist, hist_first=-1, hist_every=1,
                                            image_scale=1e-3, grad=grad2, workspace=ws2)),
                     (lam2, grad2)))
    batched = BatchedAdjoint([r[0] for r in runs]).prepare()
    batched.launch()
    torch.cuda.synchronize()
    for (_, (lam2, grad2)), (lam, grad) in zip(runs, refs):
        assert torch.equal(lam2[0], lam[0]) and torch.equal(lam2[1], lam[1])
        assert float(grad.abs().max()) > 0.0
        assert torch.equal(grad2, grad)


def test_batched_reconstructing_adjoint_equals_separate_runs():
    """Adjoint runs that rebuild the forward fields (the forward runs' final
    pairs, rings and sources), three problems of different row widths in one
    batched run (one batched launch per step count and row width), against
    separate runs: adjoint states and gradients bit for bit; the batched
    forward run that writes the rings, against separate forward runs."""
    print("batched reconstructing adjoint, 3 problems", flush=True)
    setups, single, batched_fwd = [], [], []
    for i, (ny, nx, n_el, N) in enumerate(SPECS):
        solver, sampler, array, rows, cols, dt = _problem(ny, nx, n_el, 40 + i)
        g = torch.Generator().manual_seed(50 + i)
        shape = (n_el, ny, nx)
        values = torch.randn(N + 2, generator=g).to(DEV)
        residual = torch.randn((N, n_el, n_el), generator=g).to(DEV) * 1e-2
        ring = torch.full((N, n_el, boundary_floats((ny, nx))), float("nan"), device=DEV)
        work = [torch.zeros(shape, device=DEV) for _ in range(6)]
        samples = torch.empty((N, n_el, n_el), device=DEV)
        forward_run(solver, n_el, rows, cols, values, N, samples=samples, sampler=sampler,
                    work=work, bnd=ring)
        ring2 = torch.full_like(ring, float("nan"))
        work2 = [torch.zeros(shape, device=DEV) for _ in range(6)]
        samples2 = torch.empty_like(samples)
        batched_fwd.append((solver, sampler, dict(
            n_shots=n_el, rows=rows, cols=cols, values=values, n_steps=N,
            init=(work2[0], work2[1]), work=work2, samples=samples2,
            final=(work2[0], work2[1]), bnd=ring2)))
        setups.append((solver, sampler, rows, cols, values, residual, ring, work, N, n_el,
                       shape, (ring2, work2, samples2, samples)))
    BatchedForward(batched_fwd).prepare().launch()
    torch.cuda.synchronize()
    runs = []
    for solver, sampler, rows, cols, values, residual, ring, work, N, n_el, shape, fw in setups:
        ring2, work2, samples2, samples = fw
        assert torch.equal(samples2, samples) and torch.equal(ring2, ring)
        assert torch.equal(work2[0], work[0]) and torch.equal(work2[1], work[1])
        out = {}
        for mode in ("single", "batch"):
            lam = [torch.zeros(shape, device=DEV) for _ in range(4)]
            rwork = [torch.empty(shape, device=DEV) for _ in range(4)]
            grad = torch.zeros(shape[1:], device=DEV)
            ws = adjoint_workspace(solver, n_el)
            out[mode] = (lam, rwork, grad, ws)
        lam, rwork, grad, ws = out["single"]
        adjoint_run(solver, sampler, n_el, residual, N, 0, state=(lam[0], lam[1]), work=lam,
                    image_scale=1e-3, grad=grad, workspace=ws, recon=(work[1], work[0]),
                    recon_work=rwork, bnd=ring, sources=(rows, cols, values, None))
        lam2, rwork2, grad2, ws2 = out["batch"]
        runs.append(((solver, sampler, dict(
            n_shots=n_el, residual=residual, step_hi=N, step_lo=0, state=(lam2[0], lam2[1]),
            work=lam2, image_scale=1e-3, grad=grad2, workspace=ws2, recon=(work[1], work[0]),
            recon_work=rwork2, bnd=ring, sources=(rows, cols, values))),
            (lam, grad, lam2, grad2)))
    BatchedAdjoint([r[0] for r in runs]).prepare().launch()
    torch.cuda.synchronize()
    for _, (lam, grad, lam2, grad2) in runs:
        assert torch.equal(lam2[0], lam[0]) and torch.equal(lam2[1], lam[1])
        assert float(grad.abs().max()) > 0.0
        assert torch.equal(grad2, grad)


def _gradient_problems(reconstruct):
    probs = []
    for i, (ny, nx, n_el, N) in enumerate(GRAD_SPECS):
        solver, sampler, array, rows, cols, dt = _problem(ny, nx, n_el, 30 + i)
        pulse = torch.as_tensor(ricker(dt, 400e3), dtype=F32)
        # observed data from a perturbed medium
        c = solver.model.c.clone() + 25.0
        c[: ny // 2] += 15.0
        truth = TorchFDTD2D(TorchScalarWave2D(c, solver.model.dx, device=DEV), dt)
        obs, _ = simulate_shots_batch(truth, sampler, pulse.to(DEV), array.rows, array.cols, N)
        probs.append(GradientProblem(solver, sampler, array, pulse, N, obs, checkpoint_every=16,
                                     reconstruct=reconstruct))
    return probs


@pytest.mark.parametrize("reconstruct", [False, True])
def test_uneven_gradients_match_single_problem_gradient(reconstruct):
    """Batched gradients against each problem's own gradient: the adjoint over
    recomputed segments (batched launches use two-step blocks, so sums differ
    in order) or the adjoint that rebuilds the forward fields (the same
    launches: the same bits)."""
    print(f"uneven gradients vs single-problem gradient, reconstruct {reconstruct}", flush=True)
    probs = _gradient_problems(reconstruct)
    ug = UnevenGradients(probs)
    ug.run()
    torch.cuda.synchronize()
    for p in probs:
        assert p.reconstruct == reconstruct
        eng = CheckpointedGradient(p.solver, p.sampler, p.S, p.N, checkpoint_every=p.K,
                                   reconstruct=reconstruct)
        samples = eng.forward(p.rows, p.cols, p.values)
        residual, misfit = residual_and_misfit(samples, p.observed, p.index, p.index)
        g = torch.zeros_like(p.grad)
        eng.backward(residual, g, p.solver.dt2)
        torch.cuda.synchronize()
        assert float(misfit) > 0.0
        assert abs(float(p.misfit) - float(misfit)) <= 1e-6 * float(misfit)
        err = float((p.grad - g).norm() / g.norm())
        print(f"  {tuple(p.solver.model.shape)} x {p.S} shots x {p.N} steps: rel {err:.1e}")
        assert err <= 1e-6
        if reconstruct:
            assert torch.equal(p.grad, g) and float(p.misfit) == float(misfit)


@pytest.mark.parametrize("reconstruct", [False, True])
def test_uneven_gradients_graph_replay_matches_eager(reconstruct):
    print(f"uneven gradients captured in a graph, reconstruct {reconstruct}", flush=True)
    probs = _gradient_problems(reconstruct)
    ug = UnevenGradients(probs)
    ug.run()
    torch.cuda.synchronize()
    eager = [p.grad.clone() for p in probs]
    misfits = [p.misfit.clone() for p in probs]
    side = torch.cuda.Stream(DEV)
    side.wait_stream(torch.cuda.current_stream(DEV))
    with torch.cuda.stream(side):  # warm-up on the capture stream
        ug.run()
    torch.cuda.current_stream(DEV).wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        ug.run()
    for rep in range(3):
        for p in probs:  # eager work between replays
            p.grad.fill_(123.0)
        graph.replay()
        torch.cuda.synchronize()
        for p, g, m in zip(probs, eager, misfits):
            assert torch.equal(p.grad, g), f"replay {rep}"
            assert torch.equal(p.misfit, m)
