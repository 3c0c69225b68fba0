"""swfd_residual_noalloc (built-in scratch, no allocation) against swfd_residual.

Both entry points run the same residual and misfit kernels; the one without a
workspace argument keeps its per-tile partial sums in a device array built
into the library. Two calls in a row, the second smaller, must each equal the
workspace entry point bit for bit (residual and misfit), so contents left in
the shared scratch by the first call cannot leak into the second.
"""

import ctypes

import numpy as np
import pytest
import torch

from soundwavesimulation_amd import _native

pytestmark = pytest.mark.gpu


def _call(lib, fn, n_steps, n_shots, n_rec, samples, observed, obs_index, mute, workspace=None):
    dev = samples.device
    residual = torch.full_like(samples, float("nan"))
    misfit = torch.full((1,), float("nan"), device=dev)
    args = [n_steps, n_shots, n_rec, _native.ptr(samples), _native.ptr(observed),
            observed.stride(0), observed.stride(1), _native.ptr(obs_index), _native.ptr(mute),
            _native.ptr(residual)]
    if workspace is not None:
        args.append(_native.ptr(workspace))
    args += [_native.ptr(misfit), _native.stream_of(dev)]
    _native.check(fn(*args))
    torch.cuda.synchronize(dev)
    return residual.cpu(), misfit.cpu()


def _case(n_steps, n_shots, n_rec, n_el, seed, dev):
    g = torch.Generator().manual_seed(seed)
    samples = torch.randn((n_steps, n_shots, n_rec), generator=g).to(dev)
    observed = torch.randn((n_el, n_rec, n_steps), generator=g).to(dev)
    obs_index = torch.as_tensor(np.sort(np.random.default_rng(seed).choice(
        n_el, n_shots, replace=False)).astype(np.int32)).to(dev)
    mute = obs_index.clone()
    mute[0] = -1  # one shot with no muted channel
    return samples, observed, obs_index, mute


@pytest.mark.parametrize("sizes", [((1500, 96, 96), (130, 5, 33)),
                                   ((6000, 128, 256), (4000, 16, 128))])
def test_noalloc_residual_matches_workspace_entry_point(sizes):
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    dev = torch.device("cuda", 0)
    lib = _native.load_library()
    for n_steps, n_shots, n_rec in sizes:
        print(f"noalloc residual: steps={n_steps} shots={n_shots} receivers={n_rec}", flush=True)
        n_el = max(n_rec, n_shots)
        samples, observed, obs_index, mute = _case(n_steps, n_shots, n_rec, n_el,
                                                   n_steps + n_shots, dev)
        tiles = lib.swfd_residual_workspace_floats(n_steps, n_shots, n_rec)
        assert tiles <= _native.RESIDUAL_SCRATCH_FLOATS
        workspace = torch.empty(tiles, device=dev)
        r_ws, j_ws = _call(lib, lib.swfd_residual, n_steps, n_shots, n_rec, samples, observed,
                           obs_index, mute, workspace)
        r_na, j_na = _call(lib, lib.swfd_residual_noalloc, n_steps, n_shots, n_rec, samples,
                           observed, obs_index, mute)
        assert torch.equal(r_na, r_ws)
        assert torch.equal(j_na, j_ws)
        # and both are the residual the definition gives (muted channel zeroed)
        obs_t = observed.cpu()[obs_index.cpu().long()].permute(2, 0, 1)
        expect = samples.cpu() - obs_t
        for s, m in enumerate(mute.cpu().tolist()):
            if m >= 0:
                expect[:, s, m] = 0.0
        assert torch.equal(r_na, expect)
        j64 = 0.5 * float((expect.double() ** 2).sum())
        assert abs(float(j_na[0]) - j64) <= 1e-5 * j64


def test_noalloc_residual_rejects_sizes_beyond_its_scratch():
    """Host-side size check (no device access before the rejection)."""
    lib = _native.load_library()
    fake = ctypes.c_void_p(256)  # never dereferenced: the size check comes first
    n_steps, n_shots, n_rec = 64 * 1000, 300, 256
    assert lib.swfd_residual_workspace_floats(n_steps, n_shots, n_rec) > \
        _native.RESIDUAL_SCRATCH_FLOATS
    rc = lib.swfd_residual_noalloc(n_steps, n_shots, n_rec, fake, fake, 1, 1, fake, fake, fake,
                                   fake, None)
    assert rc == _native.SWFD_EINVAL
    assert "built-in scratch" in lib.swfd_last_error().decode()
