"""The launch-shape autotuner times its candidates on the device and installs
the fastest: the tuned runs must give the same fields as untuned ones (every
form computes the same bits)."""

import numpy as np
import pytest
import torch

from soundwavesimulation_amd import RingArray, axis_centers, stable_dt
from soundwavesimulation_amd import _native
from soundwavesimulation_amd.autotune import tune
from soundwavesimulation_amd.torch_backend import TorchFDTD2D, TorchScalarWave2D, forward_run

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def test_tuned_forward_equals_default():
    ny, nx, S, N = 96, 80, 4, 13
    print(f"autotune {ny}x{nx} x {S} shots", flush=True)
    lib = _native.lib()
    lib.swfd_tune_clear()
    sp = 0.5e-3
    c = 1450.0 + 150.0 * np.random.default_rng(1).random((ny, nx))
    solver = TorchFDTD2D(TorchScalarWave2D(c, sp, device=DEV), stable_dt(1650.0, sp, sp, 0.3))
    x, y = axis_centers(nx, sp), axis_centers(ny, sp)
    array = RingArray(S, 0.4 * min(nx, ny) * sp, x, y)
    rows = torch.as_tensor(array.rows.astype(np.int32), device=DEV)
    cols = torch.as_tensor(array.cols.astype(np.int32), device=DEV)
    values = torch.randn(N + 2, generator=torch.Generator().manual_seed(3)).to(DEV)

    def run():
        work = [torch.zeros((S, ny, nx), device=DEV) for _ in range(4)]
        u_old, u = forward_run(solver, S, rows, cols, values, N, work=work)
        return u_old.clone(), u.clone()

    ref = run()
    try:
        cfg = tune(ny, nx, S, DEV, use_cache=False)
        assert set(cfg) >= {"depth1", "forward", "adjoint"}
        got = run()
    finally:
        lib.swfd_tune_clear()
    assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])
