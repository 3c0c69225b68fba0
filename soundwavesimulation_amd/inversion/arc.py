"""Arc acquisition: separate transmit and receive arcs (``invert_arc.py:294-337``).

``ArcLeastSquaresFWI`` is ``TorchLeastSquaresFWI`` with the transmit arc as
``array`` (source nodes) and a sampler of the receive arc; no transmitter is
a receiver, so no channel is muted (the reference's ``_residual`` /
``_residual_batch`` overrides keep the full residual). ``simulate_observed``
returns the ``(n_tx, n_rx, n_steps)`` gathers of every transmit element, as
the reference's helper of that name does.
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


def _batch_state(batched_solvers, solver, size):
    """The cached ``TorchFDTD2DBatch`` of ``size`` shots (created on first use)."""
    state = batched_solvers.get(size)
    if state is None:
        state = batched_solvers[size] = TorchFDTD2DBatch(solver, size)
    return state


def simulate_observed(solver, sampler, tx_array, pulse, source_values, n_steps, parallel_shots,
                      batched_solvers):
    """Fire every transmit element once; return (n_tx, n_rx, n_steps) traces.

    The transmitters are split into the fewest batches of at most
    ``parallel_shots`` elements, of near-equal size (the traces do not depend on
    the split); batch states are cached in ``batched_solvers`` by size.
    """
    n_tx = tx_array.n_elements
    per_batch = max(1, min(int(parallel_shots), n_tx))
    groups = np.array_split(np.arange(n_tx), -(-n_tx // per_batch))
    gathers = []
    for group in groups:
        traces, _ = simulate_shots_batch(solver, sampler, pulse, tx_array.rows[group],
                                         tx_array.cols[group], n_steps,
                                         batched=_batch_state(batched_solvers, solver, group.size),
                                         source_values=source_values)
        gathers.append(traces)
    return torch.cat(gathers, dim=0)
