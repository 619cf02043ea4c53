"""Match-cut insertion (ParSGA's cut-graph strategy) as the plain path's
second exact closure.

CPU: the oracle's serial restatement of ParSGA's cut-graph insertion
(oracle.cut_last_row: row i relaxes only the edges into chars labelled
neither N nor r[i], src/ParSGAoverrides/preprocessing.hpp:122-126, in
topological order, csr_align.hpp:436-447) is pinned by a hand-worked row and
against the exact closure O1: equal wherever the cut is exact (match <=
subst, match <= del + ins, ins >= subst - match), unequal on some graph for
costs outside those conditions, and unequal under mismatchN on a char
labelled N (ParSGA drops the edges into N for every base, where mismatchN's
sigma gives subst; the engine keeps them).  The engine refuses graphs and
costs the cut does not fit before it touches a device.

GPU: contexts created with CHSGA_CREATE_CUT_CLOSURE give O1's scores and
whole last rows on the plain path (int16 and int32 rows), the same with the
cut switched off (the CH sweep), valid traceback paths at those scores, and
fewer launches than the CH sweep where every read char has a cut class.
"""
import numpy as np
import pytest

from chsga_amd import ALPHABET_MATCH_N, ALPHABET_MISMATCH_N, Dataset, Engine, ReadSet, WORKLOADS, small_workload
from chsga_amd import native as N
from oracle import oracle as O
from tests import refmodels as R


def _load(tmp_path, strings, edges, name):
    p = R.write_dataset(tmp_path, strings, edges, name)
    return Dataset.load(p["char.adj"], p["str.adj"], None, p["labels"], p["label_offsets"])


def test_cut_oracle_hand_worked(tmp_path):
    """One string CAN, read CA.  Row 0 (C) is sigma(L, C); row 1 (A): the
    A at v1 matches from v0 (D = H[v0] + 0 = 0) and the exact closure gives
    v2 = min(D[v2], C[v1] + 1) = 1.  Under mismatchN, D[v2] = min(H[v2] + 1,
    H[v1] + sigma(N, A)) = 2 and ParSGA's cut drops the edge v1 -> v2 (v2 is
    labelled N), so the cut row keeps 2 there; under matchN D[v2] = 1 and
    both rows agree."""
    ds = _load(tmp_path, [b"CAN"], [], "can")
    d5 = ds.desc(ALPHABET_MISMATCH_N)
    assert list(O.exact_last_row(d5, b"CA")) == [1, 0, 1]
    assert list(O.cut_last_row(d5, b"CA")) == [1, 0, 2]
    d4 = ds.desc(ALPHABET_MATCH_N)
    assert list(O.exact_last_row(d4, b"CA")) == [1, 0, 1]
    assert list(O.cut_last_row(d4, b"CA")) == [1, 0, 1]


@pytest.mark.parametrize("costs,exact", [((0, 1, 1, 1), True), ((0, 1, 2, 3), True), ((1, 2, 1, 1), True),
                                         ((0, 3, 3, 1), False), ((0, 2, 1, 1), False)])
def test_cut_oracle_against_exact_closure(tmp_path, costs, exact):
    """Random string DAGs (one-char and longer strings, ACGT labels), reads
    over ACGTN, both alphabets: the cut rows equal the exact rows on every
    case when the costs meet the conditions and differ on some case when
    they do not (ins below subst - match)."""
    rng = np.random.default_rng(11)
    diff = total = 0
    for t in range(24):
        strings, edges = R.random_string_dag(rng, 12, 4 if t % 2 else 1, p_edge=0.2)
        ds = _load(tmp_path, strings, edges, f"r{t}")
        for a in (ALPHABET_MISMATCH_N, ALPHABET_MATCH_N):
            d = ds.desc(a)
            d.match_score, d.subst_score, d.del_score, d.ins_score = costs
            for _ in range(6):
                r = R.rbytes(rng, int(rng.integers(1, 12)), b"ACGTN")
                total += 1
                diff += int((O.exact_last_row(d, r) != O.cut_last_row(d, r)).any())
    assert total == 288
    assert (diff == 0) if exact else (diff > 0), diff


def test_cut_oracle_with_n_labels_under_match_n(tmp_path):
    """Under matchN an N label matches every read char, so ParSGA's cut
    (which drops the edges into N for every base) is exact there too."""
    rng = np.random.default_rng(12)
    for t in range(16):
        strings, edges = R.random_string_dag(rng, 10, 5 if t % 2 else 1, p_edge=0.25, alphabet=b"ACGTN")
        ds = _load(tmp_path, strings, edges, f"n{t}")
        d = ds.desc(ALPHABET_MATCH_N)  # points into ds: keep ds alive
        for _ in range(6):
            r = R.rbytes(rng, int(rng.integers(1, 14)), b"ACGTN")
            assert np.array_equal(O.exact_last_row(d, r), O.cut_last_row(d, r))


def test_cut_context_refusals(tmp_path):
    """The cut closure needs an acyclic string graph and costs where the
    cut is exact; both are refused on the host, before any device call."""
    cyc = _load(tmp_path, [b"CAT", b"GG"], [(0, 1), (1, 0)], "cyc")
    with pytest.raises(N.NativeError, match="acyclic"):
        Engine(cyc, cut=True)
    dag = _load(tmp_path, [b"CAT", b"GG"], [(0, 1)], "dag")
    d = dag.desc()
    d.subst_score, d.del_score = 3, 3
    with pytest.raises(N.NativeError, match="ins >= subst - match"):
        Engine(dag, desc=d, cut=True)


def _gpu_cases(tmp_path):
    rng = np.random.default_rng(23)
    cases = []
    for t in range(3):
        strings, edges = R.random_string_dag(rng, 40, 6 if t else 1, p_edge=0.1,
                                             alphabet=b"ACGTN" if t == 2 else b"ACGT")
        ds = _load(tmp_path, strings, edges, f"g{t}")
        cases.append((f"dag{t}", ds, ReadSet.from_list([R.rbytes(rng, int(rng.integers(1, 40)), b"ACGTN")
                                                          for _ in range(60)])))
    # a chain of 100 one-char strings labelled A: the C class is one
    # component of 99 strings (over the limit: the CH sweep), A has none
    chain = _load(tmp_path, [b"A"] * 100, [(i, i + 1) for i in range(99)], "chain")
    cases.append(("chain", chain, ReadSet.from_list([b"CCAAC", b"AAAAAAAA", b"ACACACAC", b"CXAAC"])))
    ds, rs = small_workload(seed=62, n_reads=30, length=150, error=0.1, n_rate=0.02)
    cases.append(("small", ds, rs))
    w = WORKLOADS["c1"]
    ds1 = w.dataset()
    r1 = w.readset(ds1, n_reads=8)
    cases.append(("c1", ds1, ReadSet.from_list([r1[i][:400] for i in range(len(r1))])))
    return cases


@pytest.mark.gpu
def test_cut_closure_equals_exact_oracle(tmp_path):
    for name, ds, rs in _gpu_cases(tmp_path):
        for a in (ALPHABET_MISMATCH_N, ALPHABET_MATCH_N):
            d = ds.desc(a)
            ref = O.exact_scores(d, rs)
            eng = Engine(ds, alphabet=a, cut=True)
            print("cut closure", name, a, len(rs), flush=True)
            assert np.array_equal(eng.align_combo(rs), ref)  # int16 rows where they fit
            launches_cut = eng.last_stats["kernel_launches"]
            for i in range(min(len(rs), 6)):
                assert np.array_equal(eng.last_row(rs[i]), O.exact_last_row(d, rs[i])), (name, a, i)
            tb = eng.align_read_with_traceback(rs[0], k=5)
            assert tb.score == ref[0] and O.validate_path(d, rs[0], tb.ops, tb.score)[0] == 0
            eng.set_option(N.OPT_CUT_CLOSURE, 0)
            assert np.array_equal(eng.align_combo(rs), ref)
            launches_ch = eng.last_stats["kernel_launches"]
            if name == "c1":
                assert launches_cut < launches_ch, (launches_cut, launches_ch)
            eng.set_option(N.OPT_CUT_CLOSURE, 1)
            assert np.array_equal(eng.score_batch(rs), ref)  # the batched path keeps the CH sweep
            eng.close()
    plain = Engine(small_workload()[0])
    with pytest.raises(N.NativeError, match="CHSGA_CREATE_CUT_CLOSURE"):
        plain.set_option(N.OPT_CUT_CLOSURE, 1)
    plain.close()
