"""MI355X-native ConSGA per-read graph-DP hot path (gfx950 HIP kernels behind a C ABI)."""
from .native import (ALPHABET_MATCH_N, ALPHABET_MISMATCH_N, OP_DELETE, OP_INSERT, OP_MATCHSUB, NativeError)
from .api import Dataset, Engine, ReadSet, TracebackResult
from .configs import WORKLOADS, small_workload

__all__ = [
    "ALPHABET_MATCH_N", "ALPHABET_MISMATCH_N", "OP_DELETE", "OP_INSERT", "OP_MATCHSUB", "NativeError",
    "Dataset", "Engine", "ReadSet", "TracebackResult", "WORKLOADS", "small_workload",
]
