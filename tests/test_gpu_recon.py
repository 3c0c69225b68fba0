"""The gradient whose adjoint launches rebuild the forward fields backwards in
time (``CheckpointedGradient(reconstruct=True)``, ``swfd_adjoint`` with
``recon_*``) on the device:

* the boundary ring the forward run keeps equals, bit for bit, the boundary
  cells of every stored forward field (full-width launches of four steps with
  footprint stores and their two- and three-step tails, single-step tails,
  column offset nx = 64 k + 1, segments holding both the top and bottom rows,
  bilinear transmitters);
* the adjoint states after the reverse pass equal those of the recomputing
  gradient bit for bit (the adjoint arithmetic is unchanged), the misfit is
  the same number, and the gradient agrees with the recomputing one to
  rel-L2 <= 2e-5 (the rebuilt fields round differently), for segment lengths
  that end in one-, two- and three-step launches;
* the gradient of the reference's 41 x 41 test problem against the
  reference's float64 gradient (rel-L2 <= 5e-4 on the interior mask, as for
  the recomputing gradient) and against the recomputing gradient (<= 2e-5).

The full-size checks (tests/test_gpu_full_size.py) run the rebuilding
gradient too: finite differences at 512^2 and 1024^2 through
TorchLeastSquaresFWI (which rebuilds by default) and the 1024^2 two-shot
gradient against the float64 oracle cell by cell in both modes.
"""

import os

import numpy as np
import pytest
import torch

from soundwavesimulation_amd import RingArray, _native, axis_centers, ricker, stable_dt
from soundwavesimulation_amd.inversion import interior_mask
from soundwavesimulation_amd.inversion.fwi import (CheckpointedGradient, TorchLeastSquaresFWI,
                                                   residual_and_misfit)
from soundwavesimulation_amd.phantoms import ellipse_phantom
from soundwavesimulation_amd.torch_backend import (TorchFDTD2D, TorchRingSampler,
                                                   TorchScalarWave2D, boundary_floats,
                                                   forward_run, simulate_shots_batch)

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def ring_problem(ny, nx, n_el, seed, h=0.5e-3, radius_frac=0.4):
    x = axis_centers(nx, h)
    y = axis_centers(ny, h)
    radius = radius_frac * min(nx, ny) * h
    dt = stable_dt(1650.0, h, h, cfl=0.3)
    array = RingArray(n_el, radius, x, y)
    c_true = ellipse_phantom(nx, ny, h, 6, seed, radius)
    return dict(x=x, y=y, h=h, dt=dt, array=array, c_true=c_true, radius=radius,
                pulse=ricker(dt, 300e3))


def boundary_of(field):
    """(S, ny, nx) -> (S, boundary_floats) in the ring's order."""
    f = field
    return torch.cat([f[:, 0, :], f[:, -1, :], f[:, 1:-1, 0], f[:, 1:-1, -1]], dim=1)


@pytest.mark.parametrize("ny,nx,n_steps,bilinear", [
    (130, 190, 203, False),   # four-step launches and a three-step tail
    (70, 129, 161, False),    # nx = 64 k + 1, single-step tail
    (40, 256, 98, True),      # segments holding rows 0 and ny - 1, two-step tail
    (256, 256, 400, True),
])
def test_ring_equals_the_boundary_of_every_forward_field(ny, nx, n_steps, bilinear):
    print(f"ring {ny}x{nx}, {n_steps} steps, bilinear {bilinear}", flush=True)
    pb = ring_problem(ny, nx, 12, 3)
    solver = TorchFDTD2D(TorchScalarWave2D(pb["c_true"], pb["h"], device=DEV), pb["dt"])
    sampler = TorchRingSampler(pb["array"], device=DEV)
    shots = np.arange(0, 12, 3)
    S = len(shots)
    if bilinear:
        rows, cols, w = pb["array"].bilinear_sources(shots)
        src_w = torch.as_tensor(w.astype(np.float32)).to(DEV)
    else:
        rows, cols = pb["array"].rows[shots], pb["array"].cols[shots]
        src_w = None
    r = torch.as_tensor(np.asarray(rows, np.int32)).to(DEV)
    c = torch.as_tensor(np.asarray(cols, np.int32)).to(DEV)
    values = torch.zeros(n_steps, device=DEV)
    p = torch.as_tensor(pb["pulse"], dtype=torch.float32)[:n_steps]
    values[:p.numel()] = p.to(DEV)
    P = boundary_floats((ny, nx))
    ring = torch.full((n_steps, S, P), float("nan"), device=DEV)
    samples = torch.empty((n_steps, S, sampler.n_elements), device=DEV)
    work = [torch.empty((S, ny, nx), device=DEV) for _ in range(6)]
    forward_run(solver, S, r, c, values, n_steps, samples=samples, sampler=sampler, work=work,
                src_weights=src_w, bnd=ring)
    hist = torch.empty((n_steps, S, ny, nx), device=DEV)
    samples2 = torch.empty_like(samples)
    forward_run(solver, S, r, c, values, n_steps, samples=samples2, sampler=sampler,
                hist=hist, hist_first=0, hist_every=1, src_weights=src_w)
    torch.cuda.synchronize()
    assert torch.equal(samples, samples2)
    for step in range(n_steps):
        assert torch.equal(ring[step], boundary_of(hist[step])), step


def _engines(pb, shots, n_steps, K, bilinear):
    solver = TorchFDTD2D(TorchScalarWave2D(np.full(pb["c_true"].shape, 1500.0), pb["h"],
                                           device=DEV), pb["dt"])
    truth = TorchFDTD2D(TorchScalarWave2D(pb["c_true"], pb["h"], device=DEV), pb["dt"])
    sampler = TorchRingSampler(pb["array"], device=DEV)
    arr = pb["array"]
    if bilinear:
        rows, cols, w = arr.bilinear_sources(shots)
        src_w = torch.as_tensor(w.astype(np.float32)).to(DEV)
    else:
        rows, cols, w = arr.rows[shots], arr.cols[shots], None
        src_w = None
    pulse_t = torch.as_tensor(pb["pulse"], dtype=torch.float32)
    observed = torch.zeros((arr.n_elements, arr.n_elements, n_steps), device=DEV)
    obs, _ = simulate_shots_batch(truth, sampler, pulse_t, rows, cols, n_steps,
                                  source_weights=w)
    observed[torch.as_tensor(shots)] = obs
    values = torch.zeros(n_steps, device=DEV)
    values[:min(n_steps, pulse_t.numel())] = pulse_t[:n_steps].to(DEV)
    r = torch.as_tensor(np.asarray(rows, np.int32)).to(DEV)
    c = torch.as_tensor(np.asarray(cols, np.int32)).to(DEV)
    idx = torch.as_tensor(np.asarray(shots, np.int32)).to(DEV)
    out = []
    for recon in (False, True):
        eng = CheckpointedGradient(solver, sampler, len(shots), n_steps, checkpoint_every=K,
                                   reconstruct=recon)
        samples = eng.forward(r, c, values, src_w)
        residual, misfit = residual_and_misfit(samples, observed, idx, idx)
        g = torch.zeros(pb["c_true"].shape, device=DEV)
        eng.backward(residual, g, solver.dt2)
        torch.cuda.synchronize()
        out.append((float(misfit.item()), g.cpu().numpy(), [t.clone() for t in eng.lam[:2]]))
        del eng
    return out


@pytest.mark.parametrize("ny,nx,n_steps,K,bilinear", [
    (130, 190, 401, 100, False),  # segments end in one-step launches
    (96, 257, 362, 60, False),    # nx = 64 k + 1; two-step tails
    (64, 320, 303, 1024, True),   # one segment, three-step tail; rows 0 and ny-1 together
    (256, 256, 1000, 1024, True),
])
def test_reconstructed_gradient_matches_the_recomputing_gradient(ny, nx, n_steps, K, bilinear):
    print(f"recon vs recompute {ny}x{nx}, {n_steps} steps, K {K}, bilinear {bilinear}",
          flush=True)
    pb = ring_problem(ny, nx, 16, 5)
    shots = np.array([0, 5, 9, 14])
    (J0, g0, lam0), (J1, g1, lam1) = _engines(pb, shots, n_steps, K, bilinear)
    assert J0 == J1
    assert torch.equal(lam0[0], lam1[0]) and torch.equal(lam0[1], lam1[1])
    mask = interior_mask(pb["x"], pb["y"], pb["radius"], 2 * pb["h"])
    e_all, e_in = rel(g1, g0), rel(g1[mask], g0[mask])
    print(f"  gradient rel-L2 vs recomputation: {e_all:.2e} (inside the ring {e_in:.2e})")
    assert e_all <= 2e-5


@pytest.mark.parametrize("ny,nx,n_steps,bilinear", [
    (320, 384, 700, False),  # six wavefronts per row (two side ones); ten segments
    (400, 257, 611, True),   # nx = 64 k + 1 (rows start one column left); bilinear sources
])
def test_lean_march_equals_the_general_march(ny, nx, n_steps, bilinear, monkeypatch):
    """Segments away from rows 0 and ny - 1 take the lean marches (static
    register rings, no row tests after the pipeline fills, the next segment's
    first rows stored with the same bits; the wavefronts holding columns 0 and
    nx - 1 with their one-sided Mur update); SWFD_ADJR_LEAN=0 sends them
    through the general march. Misfit, final adjoint states and gradient must
    be the same bits, with transmitters (point and bilinear) inside such
    segments and receiver footprints crossing interior and side wavefronts."""
    print(f"lean vs general march {ny}x{nx}, {n_steps} steps, bilinear {bilinear}",
          flush=True)
    pb = ring_problem(ny, nx, 16, 7)
    shots = np.arange(16)
    arr = pb["array"]
    solver = TorchFDTD2D(TorchScalarWave2D(np.full(pb["c_true"].shape, 1500.0), pb["h"],
                                           device=DEV), pb["dt"])
    truth = TorchFDTD2D(TorchScalarWave2D(pb["c_true"], pb["h"], device=DEV), pb["dt"])
    sampler = TorchRingSampler(arr, device=DEV)
    if bilinear:
        rows, cols, w = arr.bilinear_sources(shots)
        src_w = torch.as_tensor(w.astype(np.float32)).to(DEV)
    else:
        rows, cols, w = arr.rows[shots], arr.cols[shots], None
        src_w = None
    pulse_t = torch.as_tensor(pb["pulse"], dtype=torch.float32)
    observed, _ = simulate_shots_batch(truth, sampler, pulse_t, rows, cols, n_steps,
                                       source_weights=w)
    values = torch.zeros(n_steps, device=DEV)
    values[:min(n_steps, pulse_t.numel())] = pulse_t[:n_steps].to(DEV)
    r = torch.as_tensor(np.asarray(rows, np.int32)).to(DEV)
    c = torch.as_tensor(np.asarray(cols, np.int32)).to(DEV)
    idx = torch.as_tensor(shots.astype(np.int32)).to(DEV)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("SWFD_ADJR_LEAN", flag)
        eng = CheckpointedGradient(solver, sampler, len(shots), n_steps, reconstruct=True)
        samples = eng.forward(r, c, values, src_w)
        residual, misfit = residual_and_misfit(samples, observed, idx, idx)
        g = torch.zeros(pb["c_true"].shape, device=DEV)
        eng.backward(residual, g, solver.dt2)
        torch.cuda.synchronize()
        out[flag] = (float(misfit.item()), g.clone(), [t.clone() for t in eng.lam[:2]])
        del eng
    (J0, g0, lam0), (J1, g1, lam1) = out["0"], out["1"]
    assert J0 == J1
    assert torch.equal(lam0[0], lam1[0]) and torch.equal(lam0[1], lam1[1])
    assert torch.equal(g0, g1)
    assert float(g1.abs().max()) > 0.0


@pytest.mark.parametrize("ny,nx,n_steps,bilinear", [
    (320, 384, 403, False),  # six wavefronts per row (two side ones), five forward segments
    (400, 257, 389, True),   # nx = 64 k + 1; bilinear sources
    (520, 1024, 301, False),  # sixteen wavefronts per row
])
def test_lean_forward_march_equals_the_general_march(ny, nx, n_steps, bilinear, monkeypatch):
    """Footprint-store forward launches away from rows 1 and ny - 1 take the
    lean forward marches; SWFD_FWD_LEAN=0 sends them through the general
    march. Receiver samples, the boundary ring of every step and the final
    pair must be the same bits (transmitters inside lean segments, receiver
    footprints crossing interior and side wavefronts)."""
    print(f"lean vs general forward {ny}x{nx}, {n_steps} steps, bilinear {bilinear}",
          flush=True)
    pb = ring_problem(ny, nx, 16, 9)
    solver = TorchFDTD2D(TorchScalarWave2D(pb["c_true"], pb["h"], device=DEV), pb["dt"])
    sampler = TorchRingSampler(pb["array"], device=DEV)
    shots = np.arange(16)
    S = len(shots)
    if bilinear:
        rows, cols, w = pb["array"].bilinear_sources(shots)
        src_w = torch.as_tensor(w.astype(np.float32)).to(DEV)
    else:
        rows, cols = pb["array"].rows[shots], pb["array"].cols[shots]
        src_w = None
    r = torch.as_tensor(np.asarray(rows, np.int32)).to(DEV)
    c = torch.as_tensor(np.asarray(cols, np.int32)).to(DEV)
    values = torch.zeros(n_steps, device=DEV)
    p = torch.as_tensor(pb["pulse"], dtype=torch.float32)[:n_steps]
    values[:p.numel()] = p.to(DEV)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("SWFD_FWD_LEAN", flag)
        ring = torch.full((n_steps, S, boundary_floats((ny, nx))), float("nan"), device=DEV)
        samples = torch.empty((n_steps, S, sampler.n_elements), device=DEV)
        work = [torch.empty((S, ny, nx), device=DEV) for _ in range(6)]
        forward_run(solver, S, r, c, values, n_steps, samples=samples, sampler=sampler,
                    work=work, src_weights=src_w, bnd=ring)
        torch.cuda.synchronize()
        out[flag] = (samples, ring, work[0].clone(), work[1].clone())
    for a, b in zip(out["0"], out["1"]):
        assert torch.equal(a, b)
    assert float(out["1"][0].abs().max()) > 0.0


def test_reference_problem_gradient_with_reconstruction():
    g64 = np.load(os.path.join(GOLDEN, "ring41_fp64.npz"))
    sp, dt = float(g64["spacing"]), float(g64["dt"])
    x = axis_centers(41, sp)
    arr = RingArray(6, float(g64["radius"]), x, x)
    mask = g64["mask"].astype(bool)
    out = {}
    for recon in (False, True):
        solver = TorchFDTD2D(TorchScalarWave2D(g64["c_true"], sp, device=DEV), dt)
        sampler = TorchRingSampler(arr, device=DEV)
        problem = TorchLeastSquaresFWI(solver, sampler, arr, g64["pulse"], int(g64["n_steps"]),
                                       6, torch.as_tensor(g64["observed"]), g64["mask"],
                                       c_min=1400.0, c_max=2000.0, c_background=1500.0,
                                       parallel_shots=6, reconstruct=recon)
        assert problem.reconstruct == recon
        out[recon] = problem.misfit_and_grad(g64["m0"])
    J, gm = out[True]
    eJ = abs(J - float(g64["J"])) / float(g64["J"])
    e64 = rel(gm[mask], g64["g_m"][mask])
    e_re = rel(gm[mask], out[False][1][mask])
    print(f"41^2 reference problem, rebuilt fields: J rel {eJ:.2e}, gradient rel-L2 vs "
          f"float64 {e64:.2e}, vs recomputation {e_re:.2e}")
    assert J == out[False][0]
    assert eJ <= 1e-5 and e64 <= 5e-4 and e_re <= 2e-5


@pytest.mark.parametrize("ny,S,rows,groups", [
    (256, 128, 32, 128), (384, 128, 32, 64), (512, 128, 32, 64), (768, 128, 32, 32),
    (1024, 256, 32, 32), (1024, 128, 32, 32),
])
def test_reconstructing_adjoint_shapes_of_the_measured_sweep(ny, S, rows, groups):
    """The reconstructing adjoint's launch shape at the sizes of the shape sweep
    (every combination of 8/16/32-row segments and 8-128 shot groups timed on
    the device) is the fastest one found there; the recomputing adjoint keeps
    at most 32 groups."""
    if torch.cuda.get_device_properties(0).multi_processor_count != 256:
        pytest.skip("the shape rule's figures are for a 256-CU MI355X")
    shape = _native.adjoint_launch_shape(ny, ny, S, True)
    print(f"adjoint shape {ny}^2 x {S}: {shape}", flush=True)
    assert shape["full_width"] == 1 and shape["steps_per_launch"] == 4
    assert (shape["segment_rows"], shape["shot_groups"]) == (rows, groups)
    assert _native.adjoint_launch_shape(ny, ny, S, False)["shot_groups"] <= 32
