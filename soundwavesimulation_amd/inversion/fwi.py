"""Least-squares FWI with the adjoint-state gradient on the device.

Drop-in counterpart of the reference ``invert_ring.py:395-701``
(``forward_store``, ``forward_store_batch``, ``TorchLeastSquaresFWI``) on the
native kernels of :mod:`soundwavesimulation_amd.torch_backend`.

The exact gradient (``store_every=1``, full-precision history, the
reference's default) is computed by :class:`CheckpointedGradient`: the
forward run keeps only the state pair before every ``checkpoint_every``-th
step; the reverse pass re-runs each segment from its checkpoint into a
segment history and runs the adjoint over it. Recomputation reproduces the
forward fields bit for bit, so the result equals the store-everything
gradient while holding ``2 N/K + K + 1`` fields instead of ``N``.
With ``reconstruct`` (the default of :class:`TorchLeastSquaresFWI` where the
grid allows full-width launches) the reverse pass instead rebuilds the
forward fields backwards in time inside the adjoint launches, from the
checkpoints and a ring of the boundary cells of every step: no recomputation
and no segment history; the rebuilt fields differ from the forward ones by
rounding (about 1e-6 relative).
The reference's approximate options (``store_every > 1``, reduced-precision
``wave_dtype``) keep its stored-history algorithm.
"""

from __future__ import annotations

import math
import os

import numpy as np
import torch

from .. import _native
from ..torch_backend import (TorchFDTD2DBatch, _history_code, _source_values_tensor,
                             adjoint_run, adjoint_workspace, auto_batch_size, boundary_floats,
                             convert_history, forward_run)
from .medium import c_from_m, m_from_c
from .misfit import tikhonov, total_variation

_F32 = torch.float32


def stored_steps(n_steps, store_every):
    """Number of wavefield snapshots kept when storing every k-th step."""
    return -(-int(n_steps) // int(store_every))


def segment_plan(n_steps, checkpoint_every):
    """``(n_segments, n_checkpoint_slots, n_history_slots)`` of the checkpointed gradient."""
    K = int(checkpoint_every)
    n_seg = -(-int(n_steps) // K)
    return n_seg, 2 * (n_seg - 1), K + 1


def default_checkpoint_every(n_steps):
    """Segment length minimising the resident fields ``2 N/K + K`` (about sqrt(2N)),
    rounded up to a multiple of four: the adjoint then runs each segment in four-step
    launches and the recomputation (K - 1 steps) ends with a three-step launch
    instead of a single step (1024^2, 6000 steps: K = 112, 1.3 % faster than 110)."""
    k = math.sqrt(2.0 * max(int(n_steps), 1))
    return max(2, int(math.ceil(k / 4.0)) * 4 if k >= 6.0 else int(round(k)))


#: float32 staging slots of a reduced-precision history (forward chunks are
#: packed from them, adjoint chunks unpacked into them)
HISTORY_STAGING_SLOTS = 4

#: steps between the checkpoints a reconstructing gradient restarts its
#: backward rebuild of the forward field from (the rebuild drifts from the
#: forward fields by rounding only, about 1e-6 relative over thousands of steps)
RECON_EVERY = 1024


def reconstruction_available(grid_shape):
    """Whether the adjoint can rebuild the forward fields on this grid
    (full-width launches), and SWFD_RECON is not 0."""
    if os.environ.get("SWFD_RECON", "1") == "0":
        return False
    return bool(_native.lib().swfd_adjoint_recon(int(grid_shape[0]), int(grid_shape[1])))


def recon_bytes_per_shot(n_steps, grid_shape, n_receivers, checkpoint_every=None):
    """Device memory one shot of a reconstructing gradient needs: 12 work fields
    (forward, adjoint, rebuilt pairs), the shot-group images (one field per shot
    at most), checkpoints, the boundary ring of every step, samples and
    residuals."""
    cells = grid_shape[0] * grid_shape[1]
    K = int(checkpoint_every or RECON_EVERY)
    n_seg = -(-int(n_steps) // K)
    ring = int(n_steps) * (2 * grid_shape[1] + 2 * (grid_shape[0] - 2))
    return (13 + 2 * (n_seg - 1)) * cells * 4 + ring * 4 + 3 * n_steps * n_receivers * 4


def gradient_bytes_per_shot(n_steps, grid_shape, n_receivers, store_every=1, wave_dtype=None,
                            dtype=_F32, checkpoint_every=None):
    """Device memory one gradient shot needs.

    ``store_every == 1`` with full-precision history: checkpointed plan,
    ``(9 + 2 (n_seg - 1) + K + 1)`` fields (4 forward and 4 adjoint work
    fields, the shot-group images -- one field per shot for 8 to 32 shots,
    fewer above, 8 fields in all below -- checkpoints and the segment history)
    plus samples and residuals.
    Otherwise the stored history in ``wave_dtype`` (2 B per cell for float16 /
    bfloat16, held in that dtype) plus 12 float32 fields (6 forward work
    fields and the staging slots, or 4 adjoint fields, the staging slots and
    the shot-group images) and four (N, R) trace arrays, as the reference
    prices it (invert_ring.py:463-473).
    """
    cells = grid_shape[0] * grid_shape[1]
    if int(store_every) == 1 and wave_dtype in (None, _F32):
        K = checkpoint_every or default_checkpoint_every(n_steps)
        _, n_ckpt, n_hist = segment_plan(n_steps, K)
        return (9 + n_ckpt + n_hist) * cells * 4 + 3 * n_steps * n_receivers * 4
    wave_bytes = torch.empty((), dtype=wave_dtype or _F32).element_size()
    return (stored_steps(n_steps, store_every) * cells * wave_bytes + 12 * cells * 4
            + 4 * n_steps * n_receivers * 4)


def observed_bytes_per_shot(n_steps, grid_shape, n_receivers, dtype=_F32):
    cells = grid_shape[0] * grid_shape[1]
    return 12 * cells * 4 + 3 * n_steps * n_receivers * 4


def _i32(values, device):
    return torch.as_tensor(np.asarray(values, dtype=np.int32)).to(device)


def residual_and_misfit(samples, observed, obs_index, mute):
    """``residual (N, S, R)`` and device scalar ``0.5 * sum(residual^2)``.

    ``samples`` (N, S, R); ``observed`` (N_el, R, N) (any strides with the
    time axis contiguous); ``obs_index`` / ``mute`` int32 (S,) device tensors.
    """
    N, S, R = samples.shape
    if observed.stride(-1) != 1:
        observed = observed.contiguous()
    if observed.shape[1] != R or observed.shape[2] != N:
        raise ValueError("observed must have shape (elements, receivers, steps).")
    dev = samples.device
    residual = torch.empty_like(samples)
    work = torch.empty(int(_native.lib().swfd_residual_workspace_floats(N, S, R)) + 1,
                       device=dev, dtype=_F32)
    misfit = torch.empty(1, device=dev, dtype=_F32)
    _native.check(_native.lib().swfd_residual(
        N, S, R, _native.ptr(samples), _native.ptr(observed), observed.stride(0),
        observed.stride(1), _native.ptr(obs_index), _native.ptr(mute), _native.ptr(residual),
        _native.ptr(work), _native.ptr(misfit), _native.stream_of(dev)))
    return residual, misfit


class CheckpointedGradient:
    """Exact forward + adjoint gradient with checkpointing and recomputation.

    Holds the device buffers for one batch size. ``forward`` runs all steps,
    records samples and keeps checkpoints; ``backward`` takes the residual
    (N, S, R) and accumulates ``dt^2 sum_n lam^n * Lap(u^(n-1))`` into a
    gradient with respect to c^2.
    """

    def __init__(self, solver, sampler, n_shots, n_steps, checkpoint_every=None,
                 reconstruct=False):
        self.solver = solver
        self.sampler = sampler
        self.n_shots = int(n_shots)
        self.n_steps = int(n_steps)
        # reconstruct: the adjoint rebuilds the forward fields backwards in time
        # from the checkpoints and a ring of the boundary cells of every step
        # (no recomputation, no field history)
        self.reconstruct = bool(reconstruct)
        if self.reconstruct and not reconstruction_available(solver.model.shape):
            raise ValueError("this grid has no reconstructing adjoint (full-width launches).")
        default = RECON_EVERY if self.reconstruct else default_checkpoint_every(n_steps)
        self.K = int(checkpoint_every or default)
        if self.K < 2:
            raise ValueError("checkpoint_every must be at least 2.")
        self.n_seg, n_ckpt, n_hist = segment_plan(self.n_steps, self.K)
        dev = solver.model.device
        shape = (self.n_shots, *solver.model.shape)
        self.work = [torch.empty(shape, device=dev, dtype=_F32) for _ in range(4)]
        self.lam = [torch.empty(shape, device=dev, dtype=_F32) for _ in range(4)]
        self.ckpt = (torch.empty((max(n_ckpt, 2), *shape), device=dev, dtype=_F32)
                     if n_ckpt > 0 else None)
        if self.reconstruct:
            self.hist = None
            self.rwork = [torch.empty(shape, device=dev, dtype=_F32) for _ in range(4)]
            self.ring = torch.empty((self.n_steps, self.n_shots,
                                     boundary_floats(solver.model.shape)), device=dev,
                                    dtype=_F32)
        else:
            self.hist = torch.empty((n_hist, *shape), device=dev, dtype=_F32)
        self.workspace = adjoint_workspace(solver, self.n_shots)
        self.samples = torch.empty((self.n_steps, self.n_shots, sampler.n_elements),
                                   device=dev, dtype=_F32)
        self.rows = None
        self.cols = None
        self.values = None
        self.src_weights = None

    def forward(self, rows_i32, cols_i32, values, src_weights=None):
        """All forward steps; samples land in ``self.samples`` (N, S, R).
        ``src_weights`` (S, 4): bilinear sources (see ``forward_run``)."""
        self.rows, self.cols, self.values = rows_i32, cols_i32, values
        self.src_weights = src_weights
        # the adjoint's fields are idle here: they keep the intermediate steps of
        # each launch for the gather (and the ring); a reconstructing run keeps
        # the final pair (fields after steps N-2, N-1) in work[0], work[1]
        forward_run(self.solver, self.n_shots, rows_i32, cols_i32, values, self.n_steps,
                    samples=self.samples, sampler=self.sampler, work=self.work + self.lam[:2],
                    ckpt=self.ckpt, ckpt_every=self.K if self.ckpt is not None else 0,
                    final=self.reconstruct, src_weights=src_weights,
                    bnd=self.ring if self.reconstruct else None)
        return self.samples

    def lap_history(self):
        """Whether the recomputation stores Laplacians of the fields (the
        adjoint then images with one product per cell; same arithmetic, so the
        same gradient). Opt-in with SWFD_LAP_HISTORY=1: on MI355X at 1024^2 the
        recomputation's two extra field writes per launch cost more than the
        adjoint saves."""
        if os.environ.get("SWFD_LAP_HISTORY", "0") != "1":
            return False
        ny, nx = self.solver.model.shape
        return bool(_native.lib().swfd_lap_history(ny, nx))

    def backward(self, residual_nsr, grad, image_scale):
        """Accumulate the imaging condition of every step into ``grad`` (ny, nx)."""
        K, N = self.K, self.n_steps
        for t in self.lam[:2]:
            t.zero_()
        state = (self.lam[0], self.lam[1])
        if self.reconstruct:
            return self._backward_recon(residual_nsr, grad, image_scale, state)
        if self.lap_history():
            return self._backward_lap(residual_nsr, grad, image_scale, state)
        for j in range(self.n_seg - 1, -1, -1):
            n0 = j * K
            n1 = min(n0 + K, N)
            # history slot 1 holds the field before step n0 (imaged by step n0);
            # the older state of the pair is read where it is
            if j == 0:
                self.hist[0].zero_()
                self.hist[1].zero_()
                init = (self.hist[0], self.hist[1])
            else:
                self.hist[1].copy_(self.ckpt[2 * (j - 1) + 1])
                init = (self.ckpt[2 * (j - 1)], self.hist[1])
            n_re = n1 - 1 - n0
            if n_re > 0:
                forward_run(self.solver, self.n_shots, self.rows, self.cols, self.values, n_re,
                            init=init, hist=self.hist,
                            hist_first=n0 - 2, hist_every=1, step_begin=n0, work=self.work,
                            final=False, src_weights=self.src_weights)
            adjoint_run(self.solver, self.sampler, self.n_shots, residual_nsr, n1, n0,
                        state=state, work=self.lam, hist=self.hist, hist_first=n0 - 2,
                        hist_every=1, image_scale=image_scale, grad=grad,
                        workspace=self.workspace)
        return grad

    def _backward_recon(self, residual_nsr, grad, image_scale, state):
        # Each segment rebuilds the forward fields backwards from the pair
        # after its last two steps: the final state for the last segment, else
        # the checkpoint before the next segment (fields after n1-2 and n1-1).
        K, N = self.K, self.n_steps
        sources = (self.rows, self.cols, self.values, self.src_weights)
        for j in range(self.n_seg - 1, -1, -1):
            n0 = j * K
            n1 = min(n0 + K, N)
            if n1 == N:
                pair = (self.work[1], self.work[0])
            else:
                c = n1 // K - 1
                pair = (self.ckpt[2 * c + 1], self.ckpt[2 * c])
            adjoint_run(self.solver, self.sampler, self.n_shots, residual_nsr, n1, n0,
                        state=state, work=self.lam, image_scale=image_scale, grad=grad,
                        workspace=self.workspace, recon=pair, recon_work=self.rwork,
                        bnd=self.ring, sources=sources)
        return grad

    def _backward_lap(self, residual_nsr, grad, image_scale, state):
        # Each segment recomputes one step more than the field history needs,
        # so that the Laplacian of its last field is taken too; the
        # checkpoints (or zeros) are read in place.
        K, N = self.K, self.n_steps
        for j in range(self.n_seg - 1, -1, -1):
            n0 = j * K
            n1 = min(n0 + K, N)
            if j == 0:
                self.work[0].zero_()
                self.work[1].zero_()
                init = (self.work[0], self.work[1])
            else:
                init = (self.ckpt[2 * (j - 1)], self.ckpt[2 * (j - 1) + 1])
            forward_run(self.solver, self.n_shots, self.rows, self.cols, self.values, n1 - n0,
                        init=init, hist=self.hist, hist_first=n0 - 2, hist_every=1,
                        step_begin=n0, work=self.work, final=False, hist_lap=True,
                        src_weights=self.src_weights)
            adjoint_run(self.solver, self.sampler, self.n_shots, residual_nsr, n1, n0,
                        state=state, work=self.lam, hist=self.hist, hist_first=n0 - 2,
                        hist_every=1, image_scale=image_scale, grad=grad,
                        workspace=self.workspace, hist_lap=True)
        return grad


def forward_store(solver, sampler, array, pulse, tx, n_steps):
    """Single shot with the full wavefield history (invert_ring.py:395-419)."""
    n_steps = int(n_steps)
    row, col = array.inject_rows_cols(tx)
    solver.reset()
    solver.set_source(row, col)
    dev = solver.model.device
    values = _source_values_tensor(pulse, n_steps, dev)
    rc = _i32([[row], [col]], dev)
    wave = torch.empty((n_steps, 1, *solver.model.shape), device=dev, dtype=_F32)
    samples = torch.empty((n_steps, 1, sampler.n_elements), device=dev, dtype=_F32)
    u_old, u = forward_run(solver, 1, rc[0], rc[1], values, n_steps, samples=samples,
                           sampler=sampler, hist=wave, hist_first=0, hist_every=1)
    solver.u_old, solver.u = u_old[0].clone(), u[0].clone()
    return samples[:, 0, :].transpose(0, 1).contiguous(), wave[:, 0]


def _history_chunks(n_steps, store_every, slots):
    """``(first_slot, n_slots, step_begin, step_end)`` of each staging chunk:
    chunk c holds the fields after steps ``step_begin, step_begin + store_every,
    ...`` (``slots`` of them at most) and the forward advances it over
    ``[step_begin, step_end)``."""
    span = slots * store_every
    for c in range(-(-n_steps // span)):
        begin = c * span
        end = min(n_steps, begin + span)
        yield c * slots, -(-(end - begin) // store_every), begin, end


def forward_store_batch(solver, sampler, array, source_values, sources, n_steps, batched,
                        store_every=1, wave_dtype=None, bilinear_source=False):
    """Batched forward keeping every ``store_every``-th field (invert_ring.py:422-455).

    Returns traces (S, R, N) and the stored wavefields (stored, S, ny, nx) in
    ``wave_dtype``. A float16 / bfloat16 history is held in that dtype: the
    run advances in chunks whose stored fields land in a few float32 staging
    slots and are packed from there (rounded to nearest even, as the
    reference's cast is).
    """
    n_steps = int(n_steps)
    store_every = int(store_every)
    dev = solver.model.device
    src_w = None
    if bilinear_source:
        rows, cols, w = array.bilinear_sources(sources)
        src_w = torch.as_tensor(w.astype(np.float32)).to(dev).contiguous()
    else:
        rows_cols = [array.inject_rows_cols(int(tx)) for tx in sources]
        rows = np.asarray([r for r, _ in rows_cols], dtype=np.int64)
        cols = np.asarray([c for _, c in rows_cols], dtype=np.int64)
    batched.set_sources(rows, cols)
    values = _source_values_tensor(None, n_steps, dev, source_values)
    S = len(sources)
    shape = (S, *solver.model.shape)
    samples = torch.empty((n_steps, S, sampler.n_elements), device=dev, dtype=_F32)
    code = _history_code(wave_dtype)
    wave = torch.empty((stored_steps(n_steps, store_every), *shape), device=dev,
                       dtype=_F32 if code is None else wave_dtype)
    if code is None:
        u_old, u = forward_run(solver, S, batched._rows_i32, batched._cols_i32, values, n_steps,
                               samples=samples, sampler=sampler, hist=wave, hist_first=0,
                               hist_every=store_every, src_weights=src_w)
    else:
        work = [torch.empty(shape, device=dev, dtype=_F32) for _ in range(6)]
        staging = torch.empty((HISTORY_STAGING_SLOTS, *shape), device=dev, dtype=_F32)
        work[0].zero_()
        work[1].zero_()
        for slot0, n_slots, begin, end in _history_chunks(n_steps, store_every,
                                                         HISTORY_STAGING_SLOTS):
            forward_run(solver, S, batched._rows_i32, batched._cols_i32, values, end - begin,
                        init=(work[0], work[1]), samples=samples, sampler=sampler,
                        hist=staging, hist_first=begin, hist_every=store_every,
                        step_begin=begin, work=work, src_weights=src_w)
            convert_history(staging[:n_slots], wave[slot0:slot0 + n_slots])
        u_old, u = work[0], work[1]
    batched.u_old, batched.u = u_old, u
    return samples.permute(1, 2, 0).contiguous(), wave


class TorchLeastSquaresFWI:
    """Least-squares FWI on the device (invert_ring.py:483-701), same constructor."""

    def __init__(self, solver, sampler, array, pulse, n_steps, n_shots, observed, mask,
                 alpha=0.0, tv=0.0, tv_eps=1.0e-9, c_min=1400.0, c_max=2000.0,
                 c_background=1500.0, shot_seed=0, parallel_shots=16, batched_solvers=None,
                 store_every=1, wave_dtype=None, checkpoint_every=None, bilinear_source=False,
                 reconstruct=None):
        self.solver = solver
        # transmitters inject on their bilinear footprints (reciprocal data);
        # the reference fires from the nearest node
        self.bilinear_source = bool(bilinear_source)
        self.model = solver.model
        self.sampler = sampler
        self.array = array
        dev = self.model.device
        self.pulse = torch.as_tensor(pulse).to(device=dev, dtype=_F32)
        self.n_steps = int(n_steps)
        self.source_values = _source_values_tensor(self.pulse, self.n_steps, dev)
        self.n_shots = int(n_shots)
        if self.n_shots < 1 or self.n_shots > array.n_elements:
            raise ValueError("n_shots must be between 1 and the number of elements.")
        if not isinstance(observed, torch.Tensor):
            observed = torch.stack([torch.as_tensor(o) for o in observed])
        if observed.shape[0] != array.n_elements:
            raise ValueError("Observed data must contain one gather per transmitter.")
        self.observed = observed.to(device=dev, dtype=_F32).contiguous()
        self.store_every = int(store_every)
        if self.store_every < 1:
            raise ValueError("store_every must be positive.")
        self.wave_dtype = _F32 if wave_dtype is None else wave_dtype
        # the exact gradient rebuilds the forward fields in the adjoint where the
        # grid allows (None: automatic), else recomputes them from checkpoints
        if reconstruct is None:
            reconstruct = self._exact() and reconstruction_available(self.model.shape)
        self.reconstruct = bool(reconstruct) and self._exact()
        default = RECON_EVERY if self.reconstruct else default_checkpoint_every(self.n_steps)
        self.checkpoint_every = int(checkpoint_every) if checkpoint_every else default
        if parallel_shots == "auto":
            if self.reconstruct:
                per_shot = recon_bytes_per_shot(self.n_steps, self.model.shape,
                                                sampler.n_elements, self.checkpoint_every)
            else:
                per_shot = gradient_bytes_per_shot(self.n_steps, self.model.shape,
                                                   sampler.n_elements, self.store_every,
                                                   self.wave_dtype, _F32, self.checkpoint_every)
            parallel_shots = auto_batch_size(dev, self.n_shots, per_shot)
        self.parallel_shots = int(parallel_shots)
        if self.parallel_shots < 1:
            raise ValueError("parallel_shots must be positive.")
        self.parallel_shots = min(self.parallel_shots, self.n_shots)
        self._batched_solvers = {} if batched_solvers is None else batched_solvers
        self._engines = {}
        self._shot_rng = np.random.default_rng(shot_seed)
        self.active_sources = None
        self.source_history = []
        self.mask = np.asarray(mask, dtype=bool)
        self.alpha = float(alpha)
        self.tv = float(tv)
        self.tv_eps = float(tv_eps)
        self.c_min = float(c_min)
        self.c_max = float(c_max)
        self.c_background = float(c_background)

    # -- medium ----------------------------------------------------------------
    def project(self, m):
        m = np.asarray(m, dtype=float)
        projected = np.clip(m, 1.0 / self.c_max**2, 1.0 / self.c_min**2).copy()
        projected[~self.mask] = m_from_c(self.c_background)
        return projected

    def apply_medium(self, m):
        m = self.project(m)
        self.solver.set_c(c_from_m(m))
        return m

    # -- shots -----------------------------------------------------------------
    @staticmethod
    def _residual(predicted, observed, tx):
        residual = (predicted - observed).clone()
        residual[tx].zero_()
        return residual

    def _sample_sources(self):
        if self.n_shots == self.array.n_elements:
            selected = np.arange(self.array.n_elements, dtype=int)
        else:
            selected = np.sort(self._shot_rng.choice(self.array.n_elements, size=self.n_shots,
                                                     replace=False))
        self.active_sources = selected
        self.source_history.append(selected.copy())
        return selected

    def _source_batches(self, sources):
        for start in range(0, len(sources), self.parallel_shots):
            yield sources[start:start + self.parallel_shots]

    def _batched_solver(self, batch_size):
        batched = self._batched_solvers.get(batch_size)
        if batched is None:
            batched = TorchFDTD2DBatch(self.solver, batch_size)
            self._batched_solvers[batch_size] = batched
        return batched

    def _engine(self, batch_size):
        engine = self._engines.get(batch_size)
        if engine is None:
            engine = CheckpointedGradient(self.solver, self.sampler, batch_size, self.n_steps,
                                          self.checkpoint_every, reconstruct=self.reconstruct)
            self._engines[batch_size] = engine
        return engine

    def _mute_indices(self, shot_batch):
        """Receiver channel muted per shot: the transmitter (ring geometry)."""
        return shot_batch

    @staticmethod
    def _residual_batch(predicted, observed, sources):
        residual = predicted - observed
        batches = torch.arange(len(sources), device=predicted.device)
        transmitters = torch.as_tensor(np.asarray(sources), device=predicted.device)
        residual[batches, transmitters] = 0.0
        return residual

    def _rows_cols(self, shot_batch):
        dev = self.model.device
        if self.bilinear_source:
            rows, cols, _ = self.array.bilinear_sources(shot_batch)
            return _i32(rows, dev), _i32(cols, dev)
        rc = [self.array.inject_rows_cols(int(tx)) for tx in shot_batch]
        return (_i32([r for r, _ in rc], dev), _i32([c for _, c in rc], dev))

    def _source_weights(self, shot_batch):
        """(S, 4) device weights of bilinear transmitters, None for point sources."""
        if not self.bilinear_source:
            return None
        _, _, w = self.array.bilinear_sources(shot_batch)
        return torch.as_tensor(w.astype(np.float32)).to(self.model.device).contiguous()

    def _exact(self):
        return self.store_every == 1 and self.wave_dtype == _F32

    # -- objective ---------------------------------------------------------------
    def misfit(self, m):
        m = self.apply_medium(m)
        sources = self.active_sources
        if sources is None:
            sources = self._sample_sources()
        dev = self.model.device
        total = torch.zeros(1, device=dev, dtype=_F32)
        for shot_batch in self._source_batches(sources):
            S = len(shot_batch)
            rows, cols = self._rows_cols(shot_batch)
            samples = torch.empty((self.n_steps, S, self.sampler.n_elements), device=dev,
                                  dtype=_F32)
            forward_run(self.solver, S, rows, cols, self.source_values, self.n_steps,
                        samples=samples, sampler=self.sampler, final=False,
                        src_weights=self._source_weights(shot_batch))
            _, part = residual_and_misfit(samples, self.observed, _i32(shot_batch, dev),
                                          _i32(self._mute_indices(shot_batch), dev))
            total = total + part
        regularization, _ = self.regularization(m)
        return float(total.item()) + regularization

    def misfit_and_grad(self, m):
        m = self.apply_medium(m)
        sources = self._sample_sources()
        dev = self.model.device
        total = torch.zeros(1, device=dev, dtype=_F32)
        g_c2 = torch.zeros(self.model.shape, device=dev, dtype=_F32)
        for shot_batch in self._source_batches(sources):
            if self._exact():
                part = self._batch_gradient_checkpointed(shot_batch, g_c2)
            else:
                part = self._batch_gradient_stored(shot_batch, g_c2)
            total = total + part
        value = float(total.item())
        regularization, g_regularization = self.regularization(m)
        gradient = -g_c2.cpu().numpy().astype(float) / (m * m)
        gradient += g_regularization
        gradient[~self.mask] = 0.0
        return value + regularization, gradient

    def _batch_gradient_checkpointed(self, shot_batch, g_c2):
        dev = self.model.device
        engine = self._engine(len(shot_batch))
        rows, cols = self._rows_cols(shot_batch)
        samples = engine.forward(rows, cols, self.source_values, self._source_weights(shot_batch))
        residual, part = residual_and_misfit(samples, self.observed, _i32(shot_batch, dev),
                                             _i32(self._mute_indices(shot_batch), dev))
        engine.backward(residual, g_c2, self.solver.dt2)
        return part

    def _batch_gradient_stored(self, shot_batch, g_c2):
        dev = self.model.device
        batched = self._batched_solver(len(shot_batch))
        predicted, wave = forward_store_batch(self.solver, self.sampler, self.array,
                                              self.source_values, shot_batch, self.n_steps,
                                              batched, self.store_every, self.wave_dtype,
                                              self.bilinear_source)
        samples = predicted.permute(2, 0, 1).contiguous()
        residual, part = residual_and_misfit(samples, self.observed, _i32(shot_batch, dev),
                                             _i32(self._mute_indices(shot_batch), dev))
        self._adjoint_accumulate_batch(batched, wave, residual, g_c2, residual_layout="nsr")
        return part

    def regularization(self, m):
        value, gradient = tikhonov(m, self.mask, self.alpha)
        tv_value, tv_gradient = total_variation(m, self.mask, self.tv, self.tv_eps)
        return value + tv_value, gradient + tv_gradient

    def _adjoint_accumulate(self, wave, residual, g_c2):
        """Unbatched adjoint of one shot through the single-step API (invert_ring.py:668-681):
        per step the bilinear scatter of ``residual[:, n]``, an adjoint step with that
        dense source and, for n > 0, ``g_c2 += dt^2 * lam * L[wave[n - 1]]`` in place."""
        self.solver.reset()
        dt2 = self.solver.dt2
        for n in range(self.n_steps - 1, -1, -1):
            source = self.sampler.record_adjoint(residual[:, n])
            if n == 0:
                self.solver.inject_field_and_step(source, adjoint=True)
            else:
                self.solver._adjoint_image_step(source, wave[n - 1], g_c2, dt2)

    def _adjoint_accumulate_batch(self, batched, wave, residual, g_c2, residual_layout="srn"):
        """Adjoint over a stored history (invert_ring.py:683-701); adds into ``g_c2``.

        ``residual`` is (S, R, N) as the reference passes it, or (N, S, R) with
        ``residual_layout="nsr"``. A float16 / bfloat16 history is unpacked a few
        stored fields at a time into float32 staging slots, newest first, and
        the adjoint runs over the steps that image with them.
        """
        S = batched.n_shots
        residual_nsr = (residual.permute(2, 0, 1).contiguous() if residual_layout == "srn"
                        else residual.contiguous())
        dev = self.model.device
        shape = (S, *self.model.shape)
        lam = [torch.zeros(shape, device=dev, dtype=_F32) for _ in range(4)]
        state = (lam[0], lam[1])
        scale = self.store_every * batched.dt2
        wave = wave.contiguous()
        if _history_code(wave.dtype) is None:
            adjoint_run(self.solver, self.sampler, S, residual_nsr, self.n_steps, 0,
                        state=state, work=lam, hist=wave, hist_first=0,
                        hist_every=self.store_every, image_scale=scale, grad=g_c2)
        else:
            staging = torch.empty((HISTORY_STAGING_SLOTS, *shape), device=dev, dtype=_F32)
            workspace = adjoint_workspace(self.solver, S)
            chunks = list(_history_chunks(self.n_steps, self.store_every,
                                          HISTORY_STAGING_SLOTS))
            for slot0, n_slots, begin, end in reversed(chunks):
                # steps begin + 1 .. end image with this chunk's fields (step n
                # images the field after step n - 1; step 0 images nothing)
                step_hi = min(self.n_steps, end + 1)
                step_lo = begin + 1 if begin > 0 else 0
                if step_hi <= step_lo:
                    continue
                convert_history(wave[slot0:slot0 + n_slots], staging[:n_slots])
                adjoint_run(self.solver, self.sampler, S, residual_nsr, step_hi, step_lo,
                            state=state, work=lam, hist=staging, hist_first=begin,
                            hist_every=self.store_every, image_scale=scale, grad=g_c2,
                            workspace=workspace)
        batched.u_old, batched.u = lam[0], lam[1]
        return g_c2
