"""Read scheduling for the batched path.

The batched path keeps SLOTS reads in flight; the device work queue hands the
next read of `order` to a slot on the tick after its read ends.  The order
decides how long slots idle at the end of a stream (the tail), so reads are
packed longest-predicted-first.  Predictions come from the learned cost model
(chsga_amd.learned) when one is given, else from the plain cost formula
rows x V (the reference computes every cell of every row, src/align.hpp:430-463),
which is also the judge of the model's error.
"""
from __future__ import annotations

import numpy as np


def plain_cost(reads, ds) -> np.ndarray:
    V = ds.info()["V"]
    return reads.lengths().astype(np.float64) * float(V)


def plan_order(reads, ds, model=None) -> np.ndarray:
    """Queue order: longest predicted cost first (LPT), ties by read index."""
    if model is not None:
        cost = np.asarray(model.predict_time(reads, ds), dtype=np.float64)
    else:
        cost = plain_cost(reads, ds)
    return np.lexsort((np.arange(len(cost)), -cost)).astype(np.int64)


def fifo_order(reads) -> np.ndarray:
    return np.arange(len(reads), dtype=np.int64)


def simulate_ticks(lengths, order, slots: int = 64) -> int:
    """Ticks the device queue needs for `order` (same rule as the engine)."""
    import heapq
    free = [0] * slots
    heapq.heapify(free)
    T = 0
    for r in order:
        t = heapq.heappop(free)
        end = t + int(lengths[r])
        T = max(T, end)
        heapq.heappush(free, end)
    return T
