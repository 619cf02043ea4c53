"""GPU traceback: checkpointed forward pass, segment recompute in HBM, walk.

Every path must be valid (each step follows a char edge or stays, the first op
is (row 0, MATCHSUB), the last consuming row is |r|-1) and rescore to the
read's score; with the reference's predecessor priority the path equals the
oracle's walk over the full score matrix.  The staged walk (used when the
final rows are not closed) is forced on in one test and must give valid
paths of the same cost.
"""
import numpy as np
import pytest

from chsga_amd import ALPHABET_MATCH_N, Dataset, Engine, ReadSet, WORKLOADS, small_workload
from chsga_amd import native as N
from oracle import oracle as O
from tests import refmodels as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 7, 500])
def test_paths_equal_oracle_walk(k):
    ds, rs = small_workload(seed=31, n_reads=6, length=120, error=0.08)
    d = ds.desc()
    eng = Engine(ds)
    scores = eng.align_combo(rs)
    for i in range(len(rs)):
        print("traceback", k, i, flush=True)
        tb = eng.align_read_with_traceback(rs[i], k=k)
        assert tb.score == scores[i]
        rc, bad = O.validate_path(d, rs[i], tb.ops, tb.score)
        assert rc == 0, (rc, bad)
        osc, ops = O.traceback(d, rs[i])
        assert osc == tb.score
        assert ops is not None and np.array_equal(ops, tb.ops)
    eng.close()


def test_short_reads_and_matchN():
    ds, rs = small_workload(seed=32, n_reads=4, length=50, n_rate=0.1)
    d = ds.desc(ALPHABET_MATCH_N)
    eng = Engine(ds, ALPHABET_MATCH_N)
    for r in [rs[0], rs[1][:1], rs[2][:2], b"NNNN"]:
        tb = eng.align_read_with_traceback(r, k=3)
        assert O.validate_path(d, r, tb.ops, tb.score)[0] == 0
        assert tb.score == eng.align_combo(ReadSet.from_list([r]))[0]
    eng.close()


def test_staged_walk_gives_valid_paths(tmp_path):
    """Walk every row through its stages (P, U, DF) with shortcut expansion."""
    rng = np.random.default_rng(8)
    checked = 0
    for t in range(6):
        strings, edges = R.random_string_dag(rng, 30, 5, p_edge=0.08)
        p = R.write_dataset(tmp_path, strings, edges, f"sw{t}")
        ds = Dataset.load(p["char.adj"], p["str.adj"], None, p["labels"], p["label_offsets"])
        d = ds.desc()
        eng = Engine(ds)
        eng.set_option(N.OPT_STAGED_WALK, 1)
        for _ in range(3):
            r = R.rbytes(rng, int(rng.integers(2, 30)))
            tb = eng.align_read_with_traceback(r, k=4)
            assert eng.last_stats["walk_fallbacks"] == len(r) - 1
            assert tb.score == eng.align_combo(ReadSet.from_list([r]))[0]
            rc, bad = O.validate_path(d, r, tb.ops, tb.score)
            assert rc == 0, (t, r, rc, bad)
            checked += 1
        eng.close()
    assert checked == 18


def test_c1_traceback_k500(c1):
    w, ds, rs = c1
    d = ds.desc(w.alphabet)
    eng = Engine(ds, w.alphabet)
    scores = eng.align_combo(rs.subset(range(4)))
    for i in range(4):
        tb = eng.align_read_with_traceback(rs[i], k=w.traceback_k)
        assert tb.score == scores[i]
        assert O.validate_path(d, rs[i], tb.ops, tb.score)[0] == 0
        assert tb.cigar()[-1] in "MDI"
    eng.close()


def test_traceback_stream_fits_the_budget_with_learned_k():
    """The learned model proposes each read's checkpoint interval from its
    predicted peak bytes; the exact formula checks it; every traceback stays
    within the budget (measured), and its path is valid and optimal."""
    from chsga_amd import learned as L
    ds, rs = small_workload(seed=41, n_reads=5, length=300, error=0.05)
    d = ds.desc()
    V = ds.info()["V"]
    eng = Engine(ds)
    scores = eng.align_combo(rs)
    budget = L.traceback_peak_bytes(300, V, 64)  # intervals above 64 do not fit a 300-char read
    model = L.load_default()
    for use_model in (False, True):
        out = eng.traceback_stream(rs, model=model if use_model else None, budget_bytes=budget)
        for i, (tb, k) in enumerate(out):
            assert k <= 64 and L.traceback_peak_bytes(len(rs[i]), V, k) <= budget
            assert tb.score == scores[i]
            assert O.validate_path(d, rs[i], tb.ops, tb.score)[0] == 0
        assert eng.last_stats["peak_device_bytes"] <= budget
    if model is not None:
        model.close()
    eng.close()


@pytest.mark.parametrize("k", [1, 7, 500])
def test_batched_traceback_equals_oracle_walk(k):
    """The batched forward pass keeps every k-th row and the last row's
    argmin; the recovered paths equal the oracle's walk over the full score
    matrix, read by read."""
    ds, rs = small_workload(seed=43, n_reads=70, length=60, error=0.08, n_rate=0.02)
    rs = ReadSet.from_list([rs[i][: 1 + (i * 11) % 60] for i in range(len(rs))])
    d = ds.desc()
    eng = Engine(ds)
    out = eng.traceback_batch(rs, k=k)
    scores = eng.align_combo(rs)
    for i, tb in enumerate(out):
        assert tb.score == scores[i]
        osc, ops = O.traceback(d, rs[i])
        assert osc == tb.score and np.array_equal(ops, tb.ops), (k, i)
    eng.close()


def test_batched_traceback_waves_groups_and_matchN():
    """A budget of a few reads' checkpoint rows forces several waves; slot
    groups, graph replay and matchN give the same paths as the per-read
    traceback."""
    from chsga_amd import learned as L
    ds, rs = small_workload(seed=44, n_reads=40, length=90, error=0.06, n_rate=0.05)
    d = ds.desc(ALPHABET_MATCH_N)
    eng = Engine(ds, ALPHABET_MATCH_N)
    V = ds.info()["V"]
    budget = 3 * L.traceback_peak_bytes(90, V, 25)
    out = eng.traceback_batch(rs, k=25, slots=96, use_graphs=True, budget_bytes=budget)
    for i, tb in enumerate(out):
        ref = eng.align_read_with_traceback(rs[i], k=25)
        assert tb.score == ref.score and np.array_equal(tb.ops, ref.ops)
        assert O.validate_path(d, rs[i], tb.ops, tb.score)[0] == 0
    with pytest.raises(N.NativeError):
        eng.traceback_batch(rs, k=25, budget_bytes=1000)
    eng.close()


def test_c1_batched_traceback(c1):
    w, ds, rs = c1
    d = ds.desc(w.alphabet)
    eng = Engine(ds, w.alphabet)
    out = eng.traceback_batch(rs, k=w.traceback_k, use_graphs=True)
    scores = eng.score_batch(rs)
    for i, tb in enumerate(out):
        assert tb.score == scores[i]
        assert O.validate_path(d, rs[i], tb.ops, tb.score)[0] == 0
    eng.close()


@pytest.mark.parametrize("band,ring", [(1, -1), (24, -1), (0, -1), (0, 0), (0, 30)])
@pytest.mark.parametrize("k", [3, 13, 64])
def test_batched_backward_equals_per_read_backward(k, band, ring):
    """Lock-step segment recompute (up to 64 reads a round, all lanes walked
    at once) gives the paths of recomputing each read's segments alone: more
    reads than slots, uneven lengths (1..200 chars), int16 and int32 rows.
    Banded rounds (band > 0: one recompute, each lane's rows kept as a window
    of chars below its walk; 1: the default 8192 chars, here the whole
    graph's width; 24: so narrow that walks leave it and their segments are
    redone on the plain path) and whole-row rounds (band 0) with ring plans by
    free memory (-1: here one ring of k rows), the smallest (0: k2 =
    ceil(sqrt(k))) and at most 30 rows (k = 64: three snapshots, k2 = 22)."""
    ds, rs = small_workload(seed=45, n_reads=150, length=200, error=0.08, n_rate=0.02)
    rs = ReadSet.from_list([rs[i][: 1 + (i * 37) % 200] for i in range(len(rs))])
    eng = Engine(ds)
    eng.set_option(N.OPT_BACKWARD_BAND, band)
    eng.set_option(N.OPT_RING_ROWS, ring)
    print("batched backward", k, ring, flush=True)
    a = eng.traceback_batch(rs, k=k)
    eng.set_option(N.OPT_BATCHED_BACKWARD, 0)
    print("per-read backward", k, flush=True)
    b = eng.traceback_batch(rs, k=k)
    eng.set_option(N.OPT_BATCHED_BACKWARD, 1)
    eng.set_option(N.OPT_NARROW, 0)
    print("batched backward, int32 rows", k, flush=True)
    c = eng.traceback_batch(rs, k=k)
    eng.set_option(N.OPT_NARROW, 1)
    print("batched backward, two slot groups and two backward workers", k, flush=True)
    d = eng.traceback_batch(rs, k=k, slots=128)
    # the read-by-read tail: never (rounds to the end), and from the first
    # round on (every read's segments on the plain path, checkpoints from the
    # batched pool, int16 and int32)
    eng.set_option(N.OPT_BACKWARD_TAIL, 0)
    print("batched backward without the read-by-read tail", k, flush=True)
    e = eng.traceback_batch(rs, k=k)
    eng.set_option(N.OPT_BACKWARD_TAIL, 10 ** 6)
    print("batched forward, every segment read by read", k, flush=True)
    f = eng.traceback_batch(rs, k=k)
    eng.set_option(N.OPT_NARROW, 0)
    g = eng.traceback_batch(rs, k=k, slots=128)
    for i in range(len(rs)):
        assert a[i].score == b[i].score == c[i].score == d[i].score
        assert np.array_equal(a[i].ops, b[i].ops) and np.array_equal(a[i].ops, c[i].ops), (k, i)
        assert np.array_equal(a[i].ops, d[i].ops), (k, i)
        for o in (e, f, g):
            assert o[i].score == a[i].score and np.array_equal(o[i].ops, a[i].ops), (k, i)
    eng.close()


def test_batched_backward_staged_walk_gives_valid_paths(tmp_path):
    """The batched backward with every row walked through its stages: lane
    rows are made contiguous for the staged walk; paths are valid and of the
    read's score (random string DAGs, first chars with many in-neighbours)."""
    rng = np.random.default_rng(9)
    for t in range(3):
        strings, edges = R.random_string_dag(rng, 30, 5, p_edge=0.08)
        p = R.write_dataset(tmp_path, strings, edges, f"bsw{t}")
        ds = Dataset.load(p["char.adj"], p["str.adj"], None, p["labels"], p["label_offsets"])
        d = ds.desc()
        eng = Engine(ds)
        reads = [R.rbytes(rng, int(rng.integers(1, 30))) for _ in range(5)]
        rs = ReadSet.from_list(reads)
        scores = eng.align_combo(rs)
        eng.set_option(N.OPT_STAGED_WALK, 1)
        print("batched staged walk", t, flush=True)
        out = eng.traceback_batch(rs, k=4)
        assert eng.last_stats["walk_fallbacks"] == sum(len(r) - 1 for r in reads)
        for i, tb in enumerate(out):
            assert tb.score == scores[i]
            rc, bad = O.validate_path(d, reads[i], tb.ops, tb.score)
            assert rc == 0, (t, i, rc, bad)
        eng.close()


@pytest.mark.parametrize("k", [7, 40])
def test_checkpoints_in_host_memory_and_two_level_segments(k):
    """The per-read traceback with checkpoint rows in pinned host memory
    (one row kept in HBM) and with the two-level segment store (snapshots
    every ceil(sqrt(k)) rows, sub-segments of that many rows) gives the same
    paths as with everything in HBM; so does the batched traceback's
    read-by-read backward with two-level segments."""
    ds, rs = small_workload(seed=46, n_reads=6, length=150, error=0.08, n_rate=0.02)
    d = ds.desc()
    eng = Engine(ds)
    ref = [eng.align_read_with_traceback(rs[i], k=k) for i in range(len(rs))]
    assert eng.last_stats["host_bytes"] == 0
    eng.set_option(N.OPT_CKPT_HBM_ROWS, 1)
    print("host checkpoints", k, flush=True)
    host = [eng.align_read_with_traceback(rs[i], k=k) for i in range(len(rs))]
    assert eng.last_stats["host_bytes"] > 0
    eng.set_option(N.OPT_CKPT_HBM_ROWS, -1)
    eng.set_option(N.OPT_SEG_ROWS, 4)
    print("two-level segments", k, flush=True)
    two = [eng.align_read_with_traceback(rs[i], k=k) for i in range(len(rs))]
    eng.set_option(N.OPT_BATCHED_BACKWARD, 0)
    print("batched forward, read-by-read two-level backward", k, flush=True)
    bat = eng.traceback_batch(rs, k=k)
    for i in range(len(rs)):
        for other in (host[i], two[i], bat[i]):
            assert other.score == ref[i].score and np.array_equal(other.ops, ref[i].ops), (k, i)
        assert O.validate_path(d, rs[i], ref[i].ops, ref[i].score)[0] == 0
    eng.close()


def test_c5_traceback_validates(c5_ds):
    """Traceback on the largest workload's graph (116.7M chars, structural
    bubbles, matchN, k = 500): the batched traceback falls back to the
    read-by-read backward (64-slot row buffers do not fit there); checkpoint
    rows in pinned host memory with the two-level segment store give the same
    path; every path validates and rescores to the batched score."""
    w, ds = c5_ds
    rs = w.readset(ds, n_reads=4)
    clean = ReadSet.generate(ds, seed=9, n_reads=1, length=300)
    reads = ReadSet.from_list([rs[0][:600], rs[1][:400], clean[0]])
    d = ds.desc(w.alphabet)
    eng = Engine(ds, w.alphabet)
    print("c5 batched traceback", reads.lengths().tolist(), flush=True)
    out = eng.traceback_batch(reads, k=500)
    sc = eng.score_batch(reads)
    for i, tb in enumerate(out):
        assert tb.score == sc[i]
        assert O.validate_path(d, reads[i], tb.ops, tb.score)[0] == 0
    assert out[2].score == 0
    eng.set_option(N.OPT_CKPT_HBM_ROWS, 0)
    eng.set_option(N.OPT_SEG_ROWS, 64)
    print("c5 per-read traceback, host checkpoints, two-level segments", flush=True)
    tb = eng.align_read_with_traceback(reads[0], k=500)
    assert eng.last_stats["host_bytes"] > 0
    assert tb.score == out[0].score and np.array_equal(tb.ops, out[0].ops)
    eng.close()


def test_exact_closure_matches_the_exact_oracle(tmp_path):
    """CHSGA_CREATE_EXACT_CLOSURE, ParSGA's semantics (a topological
    insertion pass per row, src/ParSGAoverrides/csr_align.hpp:316-327):
    batched and plain scores equal the exact oracle (O1) on random string
    DAGs of long and one-char strings, a generated graph and c1; traceback
    paths validate at those scores; whole last rows (plain path) equal the
    oracle's exact rows and differ from ConSGA's on some read."""
    rng = np.random.default_rng(21)
    cases = []
    for t in range(3):
        strings, edges = R.random_string_dag(rng, 40, 6 if t else 1, p_edge=0.1)
        p = R.write_dataset(tmp_path, strings, edges, f"ex{t}")
        ds = Dataset.load(p["char.adj"], p["str.adj"], None, p["labels"], p["label_offsets"])
        cases.append((ds, ReadSet.from_list([R.rbytes(rng, int(rng.integers(1, 40))) for _ in range(70)])))
    cases.append(small_workload(seed=61, n_reads=40, length=120, error=0.1, n_rate=0.02))
    w = WORKLOADS["c1"]
    ds1 = w.dataset()
    r1 = w.readset(ds1, n_reads=12)
    cases.append((ds1, ReadSet.from_list([r1[i][:300] for i in range(len(r1))])))
    rows_differ = 0
    for ds, rs in cases:
        d = ds.desc()
        ref = O.exact_scores(d, rs)
        for packed in (True, False):
            eng = Engine(ds, exact=True, packed=packed)
            print("exact closure", ds.info()["n_s"], packed, flush=True)
            assert np.array_equal(eng.score_batch(rs, use_graphs=True), ref)
            assert np.array_equal(eng.align_combo(rs), ref)
            eng.set_option(N.OPT_TOP_CLOSURE, 0)
            assert np.array_equal(eng.score_batch(rs), ref)
            eng.set_option(N.OPT_TOP_CLOSURE, 1)
            out = eng.traceback_batch(rs, k=16)
            assert eng.last_stats["walk_fallbacks"] == 0
            for i in range(len(rs)):
                assert out[i].score == ref[i]
                assert O.validate_path(d, rs[i], out[i].ops, out[i].score)[0] == 0, i
            tb = eng.align_read_with_traceback(rs[0], k=5)
            assert tb.score == ref[0] and O.validate_path(d, rs[0], tb.ops, tb.score)[0] == 0
            with pytest.raises(N.NativeError):
                eng.set_option(N.OPT_STAGED_WALK, 1)
            if packed:
                for i in range(min(len(rs), 8)):
                    row = eng.last_row(rs[i])
                    assert np.array_equal(row, O.exact_last_row(d, rs[i])), i
                    rows_differ += int((row != O.consga_last_row(d, rs[i])).any())
            eng.close()
    assert rows_differ > 0


@pytest.mark.parametrize("packed", [True, False])
def test_forward_tail_on_the_plain_path(packed):
    """Once every read has started and few slots stay busy, the batched
    passes finish the running reads one by one on the plain path: scores
    equal the oracle's and every-tick batching, and the batched traceback's
    checkpoint rows (stored from the plain path for those reads) give the
    same paths, with int16 and int32 rows; in the backward the two long
    reads leave the rounds and finish read by read (against rounds to the
    end)."""
    ds, rs = small_workload(seed=47, n_reads=70, length=1400, error=0.06, n_rate=0.02)
    # two long reads and many short ones: the tail holds the long ones
    reads = ReadSet.from_list([rs[0], rs[1][:1100]] + [rs[i][: 20 + (i * 37) % 300] for i in range(2, 70)])
    ref = O.consga_scores(ds.desc(), reads)
    eng = Engine(ds, packed=packed)
    out = {}
    for tail in (0, 3, 64):
        eng.set_option(N.OPT_FORWARD_TAIL, tail)
        # the backward: rounds to the end against the default, where the two
        # long reads leave the rounds read by read
        eng.set_option(N.OPT_BACKWARD_TAIL, 0 if tail == 0 else 6)
        print("forward tail", packed, tail, flush=True)
        sc = eng.score_batch(reads, use_graphs=True)
        assert np.array_equal(sc, ref), tail
        ticks = eng.last_stats["rows"]
        for narrow in (1, 0):
            eng.set_option(N.OPT_NARROW, narrow)
            out[tail, narrow] = eng.traceback_batch(reads, k=97, use_graphs=True)
        eng.set_option(N.OPT_NARROW, 1)
        if tail == 64:
            assert ticks > 0
    d = ds.desc()
    base = out[0, 1]
    for key, o in out.items():
        for i in range(len(reads)):
            assert o[i].score == ref[i] and np.array_equal(o[i].ops, base[i].ops), (key, i)
    for i in (0, 1, 5):
        assert O.validate_path(d, reads[i], base[i].ops, base[i].score)[0] == 0
    eng.close()
