"""chsga_score_reads_noalloc (no device allocation per call, scratch in a
static device array) against chsga_score_reads, bit for bit, on two calls in
a row, the second with a smaller input; sizes beyond its maximum are
rejected."""
import ctypes as C

import numpy as np
import pytest

from chsga_amd import Engine, ReadSet, small_workload
from chsga_amd import native as N

pytestmark = pytest.mark.gpu


def _noalloc(eng, rs):
    buf, arr, lens = rs.ptrs()
    out = np.full(len(rs), -7, np.int32)
    rc = eng._lib.chsga_score_reads_noalloc(eng._h, arr, lens.ctypes.data_as(N.i64p), len(rs),
                                            out.ctypes.data_as(N.i32p))
    return rc, out


def test_noalloc_scores_equal_score_reads():
    ds, rs = small_workload(seed=41, n_reads=24, length=150, error=0.06, n_rate=0.02)
    eng = Engine(ds)
    ref = eng.align_combo(rs)
    print("no-alloc call 1:", len(rs), "reads", flush=True)
    rc, got = _noalloc(eng, rs)
    assert rc == 0, eng._lib.chsga_last_error()
    assert np.array_equal(got, ref)
    small = ReadSet.from_list([rs[i][: 1 + (i * 13) % 90] for i in range(5)])
    ref2 = eng.align_combo(small)
    print("no-alloc call 2:", len(small), "reads", flush=True)
    rc, got2 = _noalloc(eng, small)
    assert rc == 0 and np.array_equal(got2, ref2)
    # a second context on the same device shares the array and still agrees
    eng2 = Engine(ds)
    rc, got3 = _noalloc(eng2, small)
    assert rc == 0 and np.array_equal(got3, ref2)
    eng2.close()
    # more reads than the array holds: rejected before any device work
    n = 65537
    lens = np.ones(n, np.int64)
    arr = (C.c_char_p * n)(*([b"A"] * n))
    out = np.zeros(n, np.int32)
    rc = eng._lib.chsga_score_reads_noalloc(eng._h, arr, lens.ctypes.data_as(N.i64p), n, out.ctypes.data_as(N.i32p))
    assert rc == -1 and b"65536" in eng._lib.chsga_last_error()
    eng.close()
