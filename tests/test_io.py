"""On-disk formats: the reference's .npz keys, round trips without pickle."""

import numpy as np

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
