"""Multi-step forward and adjoint launches against single-step launches.

``swfd_forward`` advances several steps per launch when it is given a fourth
work buffer: two per launch in 64-column blocks, or two to four per launch in
full-width workgroups (two when receiver samples are taken). These kernels
evaluate the same float32 expressions in the same order as the single-step
kernel (stencil, point source, Mur edges and corners), so samples, history,
checkpoints and the final state must be identical (``torch.equal``), not
merely close. The shapes exercise column blocks that
end on the right boundary (nx = 63, 125), row segments with a merged tail,
grids smaller than one segment, sources on and next to the boundary, odd step
counts (a trailing single step) and a non-zero first step.
"""

import numpy as np
import pytest
import torch

from soundwavesimulation_amd import RingArray, axis_centers, stable_dt
from soundwavesimulation_amd.torch_backend import (TorchFDTD2D, TorchRingSampler,
                                                   TorchScalarWave2D, adjoint_run, forward_run)

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
F32 = torch.float32


def _medium(ny, nx, seed, spacing=1e-3):
    rng = np.random.default_rng(seed)
    c = 1450.0 + 150.0 * rng.random((ny, nx))
    dt = stable_dt(float(c.max()), spacing, spacing, cfl=0.3)
    solver = TorchFDTD2D(TorchScalarWave2D(c, spacing, device=DEV), dt)
    return solver, spacing


def _run(solver, sampler, rows, cols, values, n_steps, n_work, *, init=None, step_begin=0,
         hist_slots=0, hist_first=0, hist_every=1, ckpt_slots=0, ckpt_every=0):
    S = rows.numel()
    shape = (S, *solver.model.shape)
    work = [torch.full(shape, float("nan"), device=DEV, dtype=F32) for _ in range(n_work)]
    samples = None
    if sampler is not None:
        samples = torch.full((step_begin + n_steps, S, sampler.n_elements), float("nan"),
                             device=DEV, dtype=F32)
    hist = (torch.zeros((hist_slots, *shape), device=DEV, dtype=F32) if hist_slots else None)
    ckpt = (torch.zeros((ckpt_slots, *shape), device=DEV, dtype=F32) if ckpt_slots else None)
    final = forward_run(solver, S, rows, cols, values, n_steps, init=init, samples=samples,
                        sampler=sampler, hist=hist, hist_first=hist_first,
                        hist_every=hist_every, ckpt=ckpt, ckpt_every=ckpt_every,
                        step_begin=step_begin, work=work, final=True)
    torch.cuda.synchronize()
    out = {"old": final[0].clone(), "cur": final[1].clone()}
    if samples is not None:
        out["samples"] = samples[step_begin:].clone()
    if hist is not None:
        out["hist"] = hist
    if ckpt is not None:
        out["ckpt"] = ckpt
    return out


def _assert_identical(a, b, label):
    for key in a:
        assert torch.equal(a[key], b[key]), (
            f"{label}: {key} differs, max |diff| "
            f"{float((a[key] - b[key]).abs().max()):.3e}")


def _pulse(n, seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(n, generator=g, dtype=F32).to(DEV)


@pytest.fixture(params=["wide4", "wide3", "wide2", "blocks"])
def form(request):
    """Launch form: full-width workgroups with 4 steps per launch (the default
    where a row fits), 3 or 2, or two-step 64-column blocks (forced through
    the tuning table for the test's grid)."""
    from soundwavesimulation_amd import _native

    lib = _native.lib()
    yield request.param
    lib.swfd_tune_clear()


def _use_form(form, ny, nx):
    from soundwavesimulation_amd import _native

    lib = _native.lib()
    lib.swfd_tune_clear()
    if form == "blocks":
        _native.check(lib.swfd_tune(2, ny, nx, 0, 0, 2))
    elif form in ("wide3", "wide2"):
        _native.check(lib.swfd_tune(2, ny, nx, 0, 1, int(form[-1])))


@pytest.fixture(params=["wide2", "wide3", "wide4", "blocks"])
def adj_form(request):
    """Adjoint launch form: full-width workgroups with 2 (the default), 3 or 4
    steps per launch, or two-step 64-column blocks."""
    from soundwavesimulation_amd import _native

    lib = _native.lib()
    yield request.param
    lib.swfd_tune_clear()


def _use_adjoint_form(form, ny, nx):
    from soundwavesimulation_amd import _native

    lib = _native.lib()
    lib.swfd_tune_clear()
    if form == "blocks":
        _native.check(lib.swfd_tune(3, ny, nx, 0, 0, 2))
    else:
        _native.check(lib.swfd_tune(3, ny, nx, 0, 1, int(form[-1])))


@pytest.mark.parametrize("ny,nx,n_steps", [
    (64, 64, 41),     # one row segment, odd step count
    (200, 63, 30),    # column W-1 on the last lane of the first block; merged tail segment
    (257, 125, 24),   # column W-1 on the last lane of the second block
    (130, 190, 19),   # boundary-adjacent sources, three column blocks
    (5, 7, 9),        # smaller than one segment and one column block
    (70, 129, 22),    # nx = 64 k + 1: the full-width rows start one column left
    (40, 1000, 12),   # sixteen wavefronts per row
])
def test_two_step_forward_is_bitwise_single_step(ny, nx, n_steps, form):
    print(f"two-step forward ({form}) {ny}x{nx}, {n_steps} steps", flush=True)
    _use_form(form, ny, nx)
    solver, sp = _medium(ny, nx, seed=ny * 1000 + nx)
    rows = [1, ny - 2, ny // 2, 0, ny - 1, 1]
    cols = [1, nx - 2, nx // 2, nx // 3, 2, nx - 1]
    rows_t = torch.tensor(rows, dtype=torch.int32, device=DEV)
    cols_t = torch.tensor(cols, dtype=torch.int32, device=DEV)
    values = _pulse(n_steps, seed=n_steps)
    sampler = None
    if min(ny, nx) >= 40:
        x, y = axis_centers(nx, sp), axis_centers(ny, sp)
        radius = 0.4 * min(ny, nx) * sp
        sampler = TorchRingSampler(RingArray(12, radius, x, y), device=DEV)
    one = _run(solver, sampler, rows_t, cols_t, values, n_steps, 3)
    two = _run(solver, sampler, rows_t, cols_t, values, n_steps, 4)
    assert bool(torch.isfinite(two["cur"]).all())
    assert float(two["cur"].abs().max()) > 0.0
    _assert_identical(one, two, f"{ny}x{nx}")
    if sampler is not None:  # samples with all levels per launch (two more fields)
        six = _run(solver, sampler, rows_t, cols_t, values, n_steps, 6)
        _assert_identical(one, six, f"{ny}x{nx} with six work fields")
    if sampler is not None:  # without samples: full-width launches take all their levels
        one = _run(solver, None, rows_t, cols_t, values, n_steps, 3)
        two = _run(solver, None, rows_t, cols_t, values, n_steps, 4)
        _assert_identical(one, two, f"{ny}x{nx} without samples")


def test_two_step_history_checkpoints_and_offset_start(form):
    ny, nx, n_steps, begin = 150, 140, 23, 4
    print(f"two-step forward ({form}) with history/checkpoints {ny}x{nx}", flush=True)
    _use_form(form, ny, nx)
    solver, sp = _medium(ny, nx, seed=7)
    x, y = axis_centers(nx, sp), axis_centers(ny, sp)
    array = RingArray(8, 0.4 * min(nx, ny) * sp, x, y)
    sampler = TorchRingSampler(array, device=DEV)
    rc = [array.inject_rows_cols(i) for i in range(3)]
    rows_t = torch.tensor([r for r, _ in rc], dtype=torch.int32, device=DEV)
    cols_t = torch.tensor([c for _, c in rc], dtype=torch.int32, device=DEV)
    values = _pulse(begin + n_steps, seed=3)
    g = torch.Generator().manual_seed(11)
    init = tuple(torch.randn((3, ny, nx), generator=g, dtype=F32).to(DEV) for _ in range(2))
    kw = dict(init=init, step_begin=begin, hist_slots=6, hist_first=begin + 3, hist_every=3,
              ckpt_slots=8, ckpt_every=5)
    one = _run(solver, sampler, rows_t, cols_t, values, n_steps, 3, **kw)
    two = _run(solver, sampler, rows_t, cols_t, values, n_steps, 4, **kw)
    assert float(two["hist"].abs().sum()) > 0.0 and float(two["ckpt"].abs().sum()) > 0.0
    _assert_identical(one, two, "history/checkpoints")
    six = _run(solver, sampler, rows_t, cols_t, values, n_steps, 6, **kw)
    _assert_identical(one, six, "history/checkpoints, six work fields")
    # without samples (all levels per launch), every step kept in the history
    kw.update(hist_slots=n_steps, hist_first=begin, hist_every=1)
    one = _run(solver, None, rows_t, cols_t, values, n_steps, 3, **kw)
    two = _run(solver, None, rows_t, cols_t, values, n_steps, 4, **kw)
    _assert_identical(one, two, "history/checkpoints without samples")


def _adjoint(solver, sampler, residual, step_hi, step_lo, n_work, *, init, hist=None,
             hist_first=0, hist_every=1):
    S = residual.shape[1]
    shape = (S, *solver.model.shape)
    state = (init[0].clone(), init[1].clone())
    work = [state[0], state[1]] + [torch.full(shape, float("nan"), device=DEV, dtype=F32)
                                   for _ in range(n_work - 2)]
    grad = torch.zeros(solver.model.shape, device=DEV, dtype=F32)
    adjoint_run(solver, sampler, S, residual, step_hi, step_lo, state=state, work=work,
                hist=hist, hist_first=hist_first, hist_every=hist_every, image_scale=1e-3,
                grad=grad if hist is not None else None)
    torch.cuda.synchronize()
    return state[0].clone(), state[1].clone(), grad


@pytest.mark.parametrize("ny,nx,n_steps,hist_every", [
    (64, 64, 12, 1),    # even step count: the last pair images the last step
    (200, 63, 9, 1),    # odd: the last step runs alone and is imaged after the loop
    (257, 125, 10, 2),  # every other step has no imaging field
    (130, 190, 7, 1),
    (5, 7, 6, 1),
    (70, 129, 11, 1),   # nx = 64 k + 1
    (40, 1000, 9, 1),   # sixteen wavefronts per row
])
def test_two_step_adjoint_matches_single_step(ny, nx, n_steps, hist_every, adj_form):
    """Adjoint state bit for bit; the gradient differs only by summation order."""
    print(f"multi-step adjoint ({adj_form}) {ny}x{nx}, {n_steps} steps, hist_every {hist_every}",
          flush=True)
    _use_adjoint_form(adj_form, ny, nx)
    solver, sp = _medium(ny, nx, seed=ny + 3 * nx)
    S = 6
    if min(ny, nx) >= 40:
        x, y = axis_centers(nx, sp), axis_centers(ny, sp)
        array = RingArray(12, 0.4 * min(ny, nx) * sp, x, y)
    else:
        x, y = axis_centers(nx, sp), axis_centers(ny, sp)
        array = RingArray(3, 0.3 * min(ny, nx) * sp, x, y)
    sampler = TorchRingSampler(array, device=DEV)
    g = torch.Generator().manual_seed(ny * nx)
    R = sampler.n_elements
    residual = torch.randn((n_steps + 2, S, R), generator=g, dtype=F32).to(DEV)
    init = tuple(torch.randn((S, ny, nx), generator=g, dtype=F32).to(DEV) for _ in range(2))
    step_lo, step_hi = 2, 2 + n_steps
    hist_first = step_lo - 1
    n_hist = (step_hi - hist_first) // hist_every + 1
    hist = torch.randn((n_hist, S, ny, nx), generator=g, dtype=F32).to(DEV)
    kw = dict(init=init, hist=hist, hist_first=hist_first, hist_every=hist_every)
    o1, u1, g1 = _adjoint(solver, sampler, residual, step_hi, step_lo, 3, **kw)
    o2, u2, g2 = _adjoint(solver, sampler, residual, step_hi, step_lo, 4, **kw)
    assert torch.equal(o1, o2) and torch.equal(u1, u2), (
        f"adjoint state differs: {float((u1 - u2).abs().max()):.3e}")
    assert float(g1.abs().max()) > 0.0
    err = float((g1 - g2).norm() / g1.norm())
    print(f"  gradient rel-L2 {err:.2e}")
    assert err <= 1e-6
    # without imaging
    o3, u3, _ = _adjoint(solver, sampler, residual, step_hi, step_lo, 4, init=init)
    o4, u4, _ = _adjoint(solver, sampler, residual, step_hi, step_lo, 3, init=init)
    assert torch.equal(o3, o4) and torch.equal(u3, u4)


def _lap64(f, sp):
    """Float64 5-point Laplacian with zero outside the grid (last two axes)."""
    f = f.astype(np.float64)
    pad = np.zeros(f.shape[:-2] + (f.shape[-2] + 2, f.shape[-1] + 2))
    pad[..., 1:-1, 1:-1] = f
    inv = 1.0 / (sp * sp)
    return inv * (pad[..., 1:-1, :-2] + pad[..., 1:-1, 2:] + pad[..., :-2, 1:-1]
                  + pad[..., 2:, 1:-1] - 4.0 * f)


@pytest.fixture(params=["wide2", "wide3", "wide4"])
def wide_form(request):
    """Full-width adjoint launch forms (Laplacian histories have no two-step
    block form: the library rejects that pairing)."""
    from soundwavesimulation_amd import _native

    yield request.param
    _native.lib().swfd_tune_clear()


@pytest.mark.parametrize("ny,nx,K", [(64, 64, 10), (70, 129, 9), (130, 190, 11), (40, 1000, 7),
                                     (5, 7, 6)])
def test_laplacian_history_matches_field_history(ny, nx, K, wide_form):
    """Laplacian histories (forward_run/adjoint_run hist_lap=True): the slots
    hold the zero-padded Laplacians of the field history, and the adjoint run
    that images with them gives the same state bits and gradient as the
    single-step adjoint imaging against the fields."""
    from soundwavesimulation_amd import _native

    adj_form = wide_form
    if not _native.lib().swfd_lap_history(ny, nx):
        pytest.skip("no full-width launches for this grid")
    print(f"Laplacian history ({adj_form}) {ny}x{nx}, segment of {K} steps", flush=True)
    _use_adjoint_form(adj_form, ny, nx)
    solver, sp = _medium(ny, nx, seed=5 * ny + nx)
    S, n0 = 5, 3
    n1 = n0 + K
    x, y = axis_centers(nx, sp), axis_centers(ny, sp)
    big = min(ny, nx) >= 40
    array = RingArray(8 if big else 3, (0.35 if big else 0.3) * min(ny, nx) * sp, x, y)
    sampler = TorchRingSampler(array, device=DEV)
    rows_t = torch.tensor([1, ny - 2, ny // 2, 0, ny - 1], dtype=torch.int32, device=DEV)
    cols_t = torch.tensor([1, nx - 2, nx // 2, nx - 1, 0], dtype=torch.int32, device=DEV)
    values = _pulse(n1 + 1, seed=K)
    g = torch.Generator().manual_seed(ny + nx + K)
    init = tuple(torch.randn((S, ny, nx), generator=g, dtype=F32).to(DEV) for _ in range(2))
    shape = (S, ny, nx)
    # field history (single steps): slot k holds the output of step n0 - 2 + k
    hist_f = torch.zeros((K + 1, *shape), device=DEV)
    hist_f[0].copy_(init[0])
    hist_f[1].copy_(init[1])
    work3 = [torch.empty(shape, device=DEV) for _ in range(3)]
    forward_run(solver, S, rows_t, cols_t, values, K - 1, init=init, hist=hist_f,
                hist_first=n0 - 2, step_begin=n0, work=work3, final=False)
    # Laplacian history: one step more, slots 1..K
    hist_l = torch.full((K + 1, *shape), float("nan"), device=DEV)
    work4 = [torch.empty(shape, device=DEV) for _ in range(4)]
    forward_run(solver, S, rows_t, cols_t, values, K, init=init, hist=hist_l,
                hist_first=n0 - 2, step_begin=n0, work=work4, final=False, hist_lap=True)
    torch.cuda.synchronize()
    ref = _lap64(hist_f[1:].cpu().numpy(), sp)
    got = hist_l[1:].cpu().numpy()
    scale = float(np.abs(ref).max())
    err = float(np.abs(got - ref).max())
    print(f"  Laplacian slots: max |diff| {err:.3e} of {scale:.3e}")
    assert np.isfinite(got).all() and err <= 1e-5 * scale
    # adjoint over [n0, n1) imaging with either history
    R = sampler.n_elements
    residual = torch.randn((n1 + 1, S, R), generator=g, dtype=F32).to(DEV)
    lam0 = tuple(torch.randn(shape, generator=g, dtype=F32).to(DEV) for _ in range(2))
    kw = dict(init=lam0, hist_first=n0 - 2)
    o1, u1, g1 = _adjoint(solver, sampler, residual, n1, n0, 3, hist=hist_f, **kw)
    st = (lam0[0].clone(), lam0[1].clone())
    work = [st[0], st[1], torch.empty(shape, device=DEV), torch.empty(shape, device=DEV)]
    g2 = torch.zeros((ny, nx), device=DEV)
    adjoint_run(solver, sampler, S, residual, n1, n0, state=st, work=work, hist=hist_l,
                hist_first=n0 - 2, image_scale=1e-3, grad=g2, hist_lap=True)
    torch.cuda.synchronize()
    assert torch.equal(st[0], o1) and torch.equal(st[1], u1)
    rel = float((g2 - g1).norm() / g1.norm())
    print(f"  gradient rel-L2 {rel:.2e}")
    assert float(g1.abs().max()) > 0.0 and rel <= 1e-6
