"""On-disk formats: the reference's .npz keys, round trips without pickle."""

import os

import numpy as np
import pytest

from soundwavesimulation_amd import RingArray, axis_centers
from soundwavesimulation_amd.io import load_model, load_traces, save_model, save_traces

TRACE_KEYS = {"traces", "time_s", "pulse", "dt", "tx", "theta", "x", "y", "radius_m",
              "torch_device", "torch_dtype", "torch_compiled", "transmitter_indices"}
MODEL_KEYS = {"c_true", "c_est", "mask", "x_m", "y_m", "element_x", "element_y",
              "ring_radius_m", "ring_center", "misfit", "optimizer"}


def test_traces_round_trip(tmp_path):
    x = axis_centers(64, 1e-3)
    array = RingArray(16, 25e-3, x, x)
    traces = np.random.default_rng(0).standard_normal((16, 100)).astype(np.float32)
    c = np.full((64, 64), 1500.0)
    p = save_traces(tmp_path / "a" / "t.npz", traces, np.arange(100) * 1e-7, array, 3,
                    np.ones(20), 1e-7, c_map=c, device="cuda:0")
    d = load_traces(p)
    assert TRACE_KEYS <= set(d)
    np.testing.assert_array_equal(d["traces"], traces)
    np.testing.assert_array_equal(d["theta"], array.theta)
    np.testing.assert_array_equal(d["sound_speed_m_s"], c)
    assert int(d["tx"]) == 3 and list(d["transmitter_indices"]) == [3]
    assert str(d["torch_device"]) == "cuda:0"


def test_model_round_trip(tmp_path):
    x = axis_centers(32, 1e-3)
    array = RingArray(8, 12e-3, x, x)
    rng = np.random.default_rng(1)
    c_true, c_est = 1500 + rng.random((2, 32, 32))
    mask = rng.random((32, 32)) > 0.5
    p = save_model(tmp_path / "m.npz", c_true, c_est, {"misfit": [3.0, 2.0, 1.5]}, x, x, array,
                   mask, "cg")
    d = load_model(p)
    assert set(d) == MODEL_KEYS
    np.testing.assert_array_equal(d["c_est"], c_est)
    np.testing.assert_array_equal(d["mask"], mask)
    np.testing.assert_array_equal(d["misfit"], [3.0, 2.0, 1.5])
    assert str(d["optimizer"]) == "cg"


def test_traces_without_a_speed_map_load_without_pickle(tmp_path):
    """``c_map=None`` (the reference's default) stores an object entry; it is
    skipped instead of failing the load."""
    x = axis_centers(32, 1e-3)
    array = RingArray(8, 12e-3, x, x)
    traces = np.ones((8, 10), dtype=np.float32)
    d = load_traces(save_traces(tmp_path / "t.npz", traces, np.arange(10) * 1e-7, array, 0,
                                np.ones(4), 1e-7))
    assert TRACE_KEYS <= set(d) and "sound_speed_m_s" not in d
    np.testing.assert_array_equal(d["traces"], traces)


REFERENCE_MODEL = "/root/reference/totalseg_s1366_final_model.npz"


@pytest.mark.skipif(not os.path.exists(REFERENCE_MODEL),
                    reason="the reference's inversion archive is not present")
def test_reads_the_reference_inversion_archive(tmp_path):
    """``totalseg_s1366_final_model.npz`` (written by the reference's
    ``save_model``, ``invert_ring.py:1262-1270``) loads without pickle with the
    keys this build writes; its element positions are the ones
    :class:`RingArray` computes for that ring, and its mask is the interior
    mask of the driver's default 2-pixel margin; a re-save reads back
    identically."""
    from soundwavesimulation_amd.inversion import interior_mask

    d = load_model(REFERENCE_MODEL)
    assert set(d) == MODEL_KEYS
    x, y = d["x_m"], d["y_m"]
    n_el = d["element_x"].size
    array = RingArray(n_el, float(d["ring_radius_m"]), x, y,
                      center=tuple(float(v) for v in d["ring_center"]))
    np.testing.assert_allclose(array.x, d["element_x"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(array.y, d["element_y"], rtol=0, atol=1e-15)
    spacing = float(x[1] - x[0])
    mask = interior_mask(x, y, float(d["ring_radius_m"]), 2 * spacing)
    assert np.array_equal(mask, d["mask"])
    assert d["c_est"].shape == d["c_true"].shape == (y.size, x.size)
    assert str(d["optimizer"]) in ("cg", "adam", "gd") and d["misfit"].ndim == 1
    p = save_model(tmp_path / "again.npz", d["c_true"], d["c_est"],
                   {"misfit": list(d["misfit"])}, x, y, array, d["mask"], str(d["optimizer"]))
    again = load_model(p)
    for k in MODEL_KEYS:
        np.testing.assert_array_equal(again[k], d[k])
