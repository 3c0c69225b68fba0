"""GPU check of the surrogate warm start for multiscale FWI, at a tiny size.

The full workload (1024^2, 256 shots, three 20-iteration stages) runs from
``python -m soundwavesimulation_amd.surrogate.warm_start``; here the same code
runs on a 128^2 FWI grid with a 64^2 pair grid (32^2 operator grid), one
stage of one iteration, to check the plumbing: pairs are generated and the
operator trains, the hook keeps the corrected start exactly when its misfit is
below the blurred start's, and the first objective of the stage is evaluated
at the start the hook returned; an operator trained on 4 mm pairs corrects
blurred starts of 2 mm problems (the transfer to a finer grid the workload
makes from 1 mm pairs to its 0.25 mm FWI grid).
"""

import numpy as np
import pytest
import torch

from soundwavesimulation_amd.experiment import stable_dt
from soundwavesimulation_amd.surrogate import warm_start as ws

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def test_warm_start_pipeline_small():
    print("warm start: 128^2 FWI grid, 16 elements, 64^2 pairs", flush=True)
    rep = ws.run("warm", DEV, n=128, spacing_m=2e-3, n_elements=16, n_steps=300,
                 freqs=(150e3,), iterations=1, n_ellipses=4, seed=3,
                 surrogate_kwargs=dict(n_pairs=12, n_test=4, epochs=2, batch=4, n_sim=64,
                                       spacing_m=4e-3, f=2, n_elements=16))
    info = rep["warm_start"]
    assert rep["surrogate"]["net_grid"] == 32 and rep["surrogate"]["train"] == 8
    assert np.isfinite(rep["surrogate"]["loss_last_epoch"])
    assert info["fallback_to_plain"] == (not info["misfit_corrected"] < info["misfit_blurred"])
    kept = info["rmse_blurred"] if info["fallback_to_plain"] else info["rmse_corrected"]
    stage = rep["stages"][0]
    assert stage["rmse"][0] == pytest.approx(kept, abs=1e-3)
    start_misfit = min(info["misfit_corrected"], info["misfit_blurred"])
    assert stage["misfit_first"] == pytest.approx(start_misfit, rel=1e-6)
    assert stage["misfit_final"] <= stage["misfit_first"]


def test_plain_run_small():
    print("plain run: 128^2 FWI grid, 16 elements", flush=True)
    rep = ws.run("plain", DEV, n=128, spacing_m=2e-3, n_elements=16, n_steps=300,
                 freqs=(150e3, 300e3), iterations=1, n_ellipses=4, seed=3)
    assert "warm_start" not in rep and len(rep["stages"]) == 2
    assert rep["stages"][0]["misfit_final"] <= rep["stages"][0]["misfit_first"]
    assert rep["rmse_final"] == pytest.approx(rep["stages"][-1]["rmse"][-1], abs=1e-3)


def test_surrogate_save_load_round_trip(tmp_path):
    print("surrogate save / load", flush=True)
    op, gen, rep = ws.train_surrogate(DEV, n_pairs=6, n_test=2, epochs=1, batch=4, n_sim=64,
                                      spacing_m=4e-3, f=2, n_elements=16,
                                      duration_s=150 * 5.14e-7)
    path = tmp_path / "s.npz"
    ws.save_surrogate(path, op, rep)
    op2, rep2 = ws.load_surrogate(path, DEV)
    assert rep2 == rep
    x = torch.randn((4, 32, 32, 4), device=DEV)
    assert torch.equal(op.forward(x).clone(), op2.forward(x).clone())


def test_surrogate_start_transfers_to_a_finer_grid():
    """The warm start applies an operator trained on coarse pairs to a finer
    FWI grid (the workload: 1 mm pairs, 0.25 mm FWI). Here: pairs on a 64^2
    grid at 4 mm (32^2 operator grid), FWI problems on 128^2 at 2 mm over the
    same 256 mm field of view. On four held-out phantoms the corrected
    start's speed RMSE inside the ring must stay below the blurred start's
    (mean over the four, and on at least three of them)."""
    print("surrogate transfer: 4 mm pairs -> 2 mm FWI grid", flush=True)
    n_steps = 300
    dt = stable_dt(1650.0, 2e-3, 2e-3, cfl=0.3)
    op, gen, rep = ws.train_surrogate(DEV, n_pairs=320, n_test=32, epochs=80, batch=32,
                                      n_sim=64, spacing_m=4e-3, f=2, n_elements=16,
                                      duration_s=n_steps * dt)
    assert rep["net_grid"] == 32
    blurred, corrected = [], []
    for seed in (41, 42, 43, 44):
        r = ws.run("warm", DEV, n=128, spacing_m=2e-3, n_elements=16, n_steps=n_steps,
                   freqs=(150e3,), iterations=1, n_ellipses=6, seed=seed, surrogate=(op, rep),
                   verbose=False)
        blurred.append(r["warm_start"]["rmse_blurred"])
        corrected.append(r["warm_start"]["rmse_corrected"])
    print(f"  operator-grid test RMSE {rep['test_rmse_m_s']}; FWI-grid start RMSE "
          f"blurred {np.round(blurred, 2)}, corrected {np.round(corrected, 2)} m/s")
    assert np.mean(corrected) < np.mean(blurred)
    assert sum(c < b for c, b in zip(corrected, blurred)) >= 3
