"""Arc acquisition: separate transmit and receive arcs (``invert_arc.py:294-337``).

``ArcLeastSquaresFWI`` is ``TorchLeastSquaresFWI`` with the transmit arc as
``array`` (source nodes) and a sampler of the receive arc; no transmitter is
a receiver, so no channel is muted (the reference's ``_residual`` /
``_residual_batch`` overrides keep the full residual). ``simulate_observed``
fires every transmit element once, in batches, and returns
``(n_tx, n_rx, n_steps)`` traces like the reference's helper.
"""

from __future__ import annotations

import numpy as np
import torch

from ..torch_backend import TorchFDTD2DBatch, simulate_shots_batch
from .fwi import TorchLeastSquaresFWI


class ArcLeastSquaresFWI(TorchLeastSquaresFWI):
    """FWI with a transmit arc (``array``) and a receive arc (``sampler``)."""

    @staticmethod
    def _residual(predicted, observed, tx):
        return predicted - observed

    @staticmethod
    def _residual_batch(predicted, observed, sources):
        return predicted - observed

    def _mute_indices(self, shot_batch):
        return np.full(len(shot_batch), -1, dtype=np.int64)


def simulate_observed(solver, sampler, tx_array, pulse, source_values, n_steps, parallel_shots,
                      batched_solvers):
    """Fire every transmit element once; return (n_tx, n_rx, n_steps) traces."""
    model = solver.model
    observed = torch.empty((tx_array.n_elements, sampler.n_elements, n_steps),
                           device=model.device, dtype=model.dtype)
    batch_size = min(parallel_shots, tx_array.n_elements)
    for start in range(0, tx_array.n_elements, batch_size):
        sources = np.arange(start, min(start + batch_size, tx_array.n_elements))
        batched = batched_solvers.get(len(sources))
        if batched is None:
            batched = TorchFDTD2DBatch(solver, len(sources))
            batched_solvers[len(sources)] = batched
        traces, _ = simulate_shots_batch(solver, sampler, pulse, tx_array.rows[sources],
                                         tx_array.cols[sources], n_steps, batched=batched,
                                         source_values=source_values)
        observed[start:start + len(sources)].copy_(traces)
    return observed
