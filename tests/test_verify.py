"""The product's path verification (chsga_host_verify_path, the drivers'
check of src/ConSGA_traceback.cpp:184-225) against the oracle's paths and
planted mistakes (no GPU)."""
import numpy as np

from chsga_amd.api import verify_path
from oracle import oracle as O


def test_verify_accepts_oracle_paths_and_rejects_mistakes(small):
    ds, rs = small
    d = ds.desc()
    for i in range(len(rs)):
        sc, ops = O.traceback(d, rs[i])
        assert verify_path(d, rs[i], ops, sc) == (0, -1)
    sc, ops = O.traceback(d, rs[0])
    assert verify_path(d, rs[0], ops, sc + 1)[0] == 6
    assert verify_path(d, rs[0], ops[:0], sc)[0] == 1
    bad = ops.copy()
    bad[0]["op"] = 2
    assert verify_path(d, rs[0], bad, sc)[0] == 2
    k = len(ops) // 2
    bad = ops.copy()
    bad[k]["vertex"] = (int(bad[k]["vertex"]) + 7) % d.V
    assert verify_path(d, rs[0], bad, sc)[0] == 3
    assert verify_path(d, rs[0], np.delete(ops, k), sc)[0] in (3, 6)
    bad = ops.copy()
    bad[k]["vertex"] = d.V
    assert verify_path(d, rs[0], bad, sc) == (4, k)
    assert verify_path(d, rs[0], ops[:-1], sc)[0] in (5, 6)
    # the verdicts agree with the oracle's validator on every mutation
    rng = np.random.default_rng(3)
    for _ in range(200):
        bad = ops.copy()
        j = int(rng.integers(len(bad)))
        f = int(rng.integers(3))
        if f == 0:
            bad[j]["op"] = int(rng.integers(3))
        elif f == 1:
            bad[j]["vertex"] = int(rng.integers(d.V))
        else:
            bad[j]["row"] = max(0, int(bad[j]["row"]) + int(rng.integers(-1, 2)))
        s = sc + int(rng.integers(-1, 2))
        assert (verify_path(d, rs[0], bad, s)[0] == 0) == (O.validate_path(d, rs[0], bad, s)[0] == 0)
