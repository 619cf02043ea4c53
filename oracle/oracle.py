"""ctypes wrapper of oracle/build/liboracle.so -- the CPU restatement of ConSGA.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker (or the timed CPU
baseline), never the thing measured on the GPU or shipped.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from chsga_amd import native as N

_LIB = None
_PATH = Path(__file__).resolve().parent / "build" / "liboracle.so"


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not _PATH.exists():
            raise RuntimeError(f"{_PATH} is not built: run `python -m chsga_amd.build`")
        L = C.CDLL(str(_PATH))
        G = C.POINTER(N.GraphDesc)
        L.oracle_threads.restype, L.oracle_threads.argtypes = C.c_int, [C.c_int]
        L.oracle_consga_scores.argtypes = [G, C.c_char_p, N.i64p, C.c_int64, N.i32p]
        L.oracle_consga_rows.argtypes = [G, C.c_char_p, C.c_int64, N.i32p]
        L.oracle_consga_last_row.argtypes = [G, C.c_char_p, C.c_int64, N.i32p]
        L.oracle_exact_last_row.argtypes = [G, C.c_char_p, C.c_int64, N.i32p]
        L.oracle_cut_last_row.argtypes = [G, C.c_char_p, C.c_int64, N.i32p]
        L.oracle_exact_scores.argtypes = [G, C.c_char_p, N.i64p, C.c_int64, N.i32p]
        L.oracle_traceback.restype = C.c_int64
        L.oracle_traceback.argtypes = [G, C.c_char_p, C.c_int64, C.POINTER(N.AlignOp), C.c_int64, N.i32p]
        L.oracle_validate_path.argtypes = [G, C.c_char_p, C.c_int64, C.POINTER(N.AlignOp), C.c_int64, C.c_int32,
                                           N.i64p]
        L.oracle_pasgal_align.argtypes = [G, C.c_char_p, N.i64p, C.c_int64, C.POINTER(N.PasgalParams),
                                          C.POINTER(N.PasgalResult)]
        L.oracle_pasgal_free.restype = None
        L.oracle_pasgal_free.argtypes = [C.POINTER(N.PasgalResult), C.c_int64]
        L.oracle_sigma.restype = C.c_int32
        L.oracle_sigma.argtypes = [C.c_int32, C.c_char, C.c_char, C.c_int32, C.c_int32]
        _LIB = L
    return _LIB


def threads(n: int = 0) -> int:
    return lib().oracle_threads(n)


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed: {rc}")


def consga_scores(desc: N.GraphDesc, reads) -> np.ndarray:
    n = len(reads)
    out = np.zeros(n, np.int32)
    off = np.ascontiguousarray(reads.offsets, np.int64)
    _ok(lib().oracle_consga_scores(C.byref(desc), reads.data, off.ctypes.data_as(N.i64p), n,
                                   out.ctypes.data_as(N.i32p)), "scores")
    return out


def exact_scores(desc: N.GraphDesc, reads) -> np.ndarray:
    n = len(reads)
    out = np.zeros(n, np.int32)
    off = np.ascontiguousarray(reads.offsets, np.int64)
    _ok(lib().oracle_exact_scores(C.byref(desc), reads.data, off.ctypes.data_as(N.i64p), n,
                                  out.ctypes.data_as(N.i32p)), "exact scores")
    return out


def consga_rows(desc: N.GraphDesc, read: bytes) -> np.ndarray:
    out = np.zeros((len(read), desc.V), np.int32)
    _ok(lib().oracle_consga_rows(C.byref(desc), read, len(read), out.ctypes.data_as(N.i32p)), "rows")
    return out


def consga_last_row(desc: N.GraphDesc, read: bytes) -> np.ndarray:
    out = np.zeros(desc.V, np.int32)
    _ok(lib().oracle_consga_last_row(C.byref(desc), read, len(read), out.ctypes.data_as(N.i32p)), "last row")
    return out


def exact_last_row(desc: N.GraphDesc, read: bytes) -> np.ndarray:
    out = np.zeros(desc.V, np.int32)
    _ok(lib().oracle_exact_last_row(C.byref(desc), read, len(read), out.ctypes.data_as(N.i32p)), "exact row")
    return out


def cut_last_row(desc: N.GraphDesc, read: bytes) -> np.ndarray:
    out = np.zeros(desc.V, np.int32)
    _ok(lib().oracle_cut_last_row(C.byref(desc), read, len(read), out.ctypes.data_as(N.i32p)), "cut row")
    return out


OP_DTYPE = np.dtype([("row", np.int64), ("vertex", np.uint32), ("op", np.uint8)], align=True)


def traceback(desc: N.GraphDesc, read: bytes):
    """(score, ops) or (score, None) when the reference walk finds no predecessor."""
    cap = 4 * len(read) + desc.V + 16
    buf = (N.AlignOp * cap)()
    sc = C.c_int32()
    n = lib().oracle_traceback(C.byref(desc), read, len(read), buf, cap, C.byref(sc))
    if n == -1:
        return int(sc.value), None
    if n < 0:
        raise RuntimeError("oracle traceback failed")
    raw = C.string_at(buf, n * C.sizeof(N.AlignOp))
    return int(sc.value), np.frombuffer(raw, dtype=OP_DTYPE).copy()


def validate_path(desc: N.GraphDesc, read: bytes, ops: np.ndarray, score: int):
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    bad = C.c_int64(-1)
    rc = lib().oracle_validate_path(C.byref(desc), read, len(read), ops.ctypes.data_as(C.POINTER(N.AlignOp)),
                                    len(ops), score, C.byref(bad))
    return rc, bad.value


def sigma(alphabet: int, a: bytes, b: bytes, match: int = 0, subst: int = 1) -> int:
    return lib().oracle_sigma(alphabet, a, b, match, subst)


def pasgal_align(desc: N.GraphDesc, reads, match: int = 1, mismatch: int = 1, ins: int = 1, dele: int = 1,
                 score_only: bool = False) -> list:
    """O5: PaSGAL_match's alignToDAGLocal, one dict per read (as
    chsga_amd.api.PasgalAligner.align)."""
    n = len(reads)
    off = np.ascontiguousarray(reads.offsets, np.int64)
    prm = N.PasgalParams(match, mismatch, ins, dele, 1 if score_only else 0)
    res = (N.PasgalResult * max(n, 1))()
    _ok(lib().oracle_pasgal_align(C.byref(desc), reads.data, off.ctypes.data_as(N.i64p), n, C.byref(prm), res),
        "pasgal")
    try:
        return [res[i].as_dict() for i in range(n)]
    finally:
        lib().oracle_pasgal_free(res, n)
