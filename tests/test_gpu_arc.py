"""Arc acquisition FWI (invert_arc.py:294-337) against the float64 oracle.

Separate transmit and receive arcs, no muted channel: the misfit and the
gradient with respect to slowness squared match the oracle (tx = -1, i.e.
nothing muted) to the float32 tolerances of the ring tests.
"""

import numpy as np
import pytest
import torch

import oracle
from soundwavesimulation_amd import axis_centers, ricker, stable_dt
from soundwavesimulation_amd.array import ArcArray
from soundwavesimulation_amd.inversion.arc import ArcLeastSquaresFWI, simulate_observed
from soundwavesimulation_amd.inversion.medium import interior_mask, m_from_c
from soundwavesimulation_amd.torch_backend import (TorchFDTD2D, TorchRingSampler,
                                                   TorchScalarWave2D, _source_values_tensor)

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def test_arc_gradient_matches_oracle():
    n, sp, N = 41, 2e-3, 260
    x = axis_centers(n, sp)
    dt = stable_dt(1650.0, sp, sp, cfl=0.3)
    pulse = ricker(dt, 60e3)
    tx = ArcArray(5, 0.030, 100.0, 180.0, x, x)
    rx = ArcArray(7, 0.030, 120.0, 0.0, x, x)
    yy, xx = np.meshgrid(x, x, indexing="ij")
    c_true = np.where(xx ** 2 + yy ** 2 < (0.012) ** 2, 1560.0, 1500.0)
    truth = TorchFDTD2D(TorchScalarWave2D(c_true, sp, device=DEV), dt)
    sampler = TorchRingSampler(rx, device=DEV)
    pulse_t = torch.as_tensor(pulse, dtype=torch.float32, device=DEV)
    values = _source_values_tensor(pulse_t, N, DEV)
    observed = simulate_observed(truth, sampler, tx, pulse_t, values, N, 8, {})
    solver = TorchFDTD2D(TorchScalarWave2D(np.full((n, n), 1500.0), sp, device=DEV), dt)
    mask = interior_mask(x, x, 0.030, 2 * sp)
    fwi = ArcLeastSquaresFWI(solver, sampler, tx, pulse_t, N, tx.n_elements, observed, mask,
                             c_min=1400.0, c_max=1650.0)
    m0 = m_from_c(np.full((n, n), 1500.0))
    J, g = fwi.misfit_and_grad(m0)
    r0, c0, w = rx.footprint()
    obs = observed.cpu().double().numpy()
    Jo, gc2, _ = oracle.gradient(np.full((n, n), 1500.0), sp, sp, dt, tx.rows, tx.cols, pulse, N,
                                 r0, c0, w, obs, np.full(tx.n_elements, -1))
    go = -gc2 / (m0 * m0)
    go[~mask] = 0.0
    eJ = abs(J - Jo) / Jo
    eg = np.linalg.norm(g[mask] - go[mask]) / np.linalg.norm(go[mask])
    print(f"arc: J rel {eJ:.2e}, g_m rel-L2 {eg:.2e}")
    assert Jo > 0.0
    assert eJ <= 1e-4
    assert eg <= 5e-4
