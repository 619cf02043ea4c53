"""GPU parity: the HIP engine against the oracle, bit-exact (int32 scores and rows).

Covers both alphabets, the plain path (one read per launch sequence) and the
batched path (slot-interleaved rows + device work queue), uneven read lengths,
one-char reads, N characters, multi-component graphs, long string vertices,
the committed golden fixtures, and at full c3 size the size-independent
properties (plain == batched, 0 <= score <= |r|, error-free reads score 0).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from chsga_amd import ALPHABET_MATCH_N, ALPHABET_MISMATCH_N, Dataset, Engine, ReadSet, WORKLOADS, small_workload
from chsga_amd import native as N
from chsga_amd.scheduler import plan_order
from oracle import oracle as O
from tests import refmodels as R

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("alphabet", [ALPHABET_MISMATCH_N, ALPHABET_MATCH_N])
def test_last_rows_match_oracle(alphabet):
    ds, rs = small_workload(seed=11, n_reads=6, length=90, n_rate=0.05)
    d = ds.desc(alphabet)
    eng = Engine(ds, alphabet)
    for i in range(len(rs)):
        _say("last row", alphabet, i, len(rs[i]))
        assert np.array_equal(eng.last_row(rs[i]), O.consga_last_row(d, rs[i]))
    eng.close()


@pytest.mark.parametrize("alphabet", [ALPHABET_MISMATCH_N, ALPHABET_MATCH_N])
def test_scores_plain_and_batched_match_oracle(alphabet):
    ds, rs = small_workload(seed=12, n_reads=90, length=140, n_rate=0.02)
    # uneven lengths including one-char reads
    reads = [rs[i][: 1 + (i * 37) % 140] for i in range(len(rs))] + [b"N" * 9, b"A", b"Z" * 3]
    rs = ReadSet.from_list(reads)
    ref = O.consga_scores(ds.desc(alphabet), rs)
    eng = Engine(ds, alphabet)
    _say("plain", alphabet)
    assert np.array_equal(eng.align_combo(rs), ref)
    _say("batched fifo")
    assert np.array_equal(eng.score_batch(rs), ref)
    _say("batched planned + graphs")
    assert np.array_equal(eng.score_batch(rs, order=plan_order(rs, ds), use_graphs=True), ref)
    _say("batched 16 slots")
    assert np.array_equal(eng.score_batch(rs, slots=16), ref)
    eng.close()


def test_batch_calls_back_to_back_and_async():
    ds, rs = small_workload(seed=13, n_reads=70, length=60)
    ref = O.consga_scores(ds.desc(), rs)
    eng = Engine(ds)
    a = eng.score_batch(rs)
    b = eng.score_batch(rs.subset(range(10)))
    assert np.array_equal(a, ref) and np.array_equal(b, ref[:10])
    eng.close()


@pytest.mark.parametrize("slots", [128, 96, 200])
def test_slot_groups_match_oracle(slots):
    """More than 64 slots run as groups of 64 on separate streams, each with
    its share of the queue; scores stay exact with and without graphs."""
    ds, rs = small_workload(seed=23, n_reads=150, length=70, error=0.05, n_rate=0.02)
    ref = O.consga_scores(ds.desc(), rs)
    eng = Engine(ds)
    _say("groups", slots)
    assert np.array_equal(eng.score_batch(rs, slots=slots), ref)
    assert np.array_equal(eng.score_batch(rs, slots=slots, order=plan_order(rs, ds), use_graphs=True), ref)
    assert np.array_equal(eng.score_batch(rs.subset(range(7)), slots=slots), ref[:7])
    eng.close()


@pytest.mark.parametrize("costs", [(0, 1, 1, 0), (0, 2, 3, 0), (1, 3, 2, 1)])
def test_other_edit_costs(costs):
    """Non-unit costs, including ins = 0, under which an insertion closure of
    row 0 (which the reference does not run) would change the scores."""
    ds, rs = small_workload(seed=29, n_reads=90, length=40, error=0.1, n_rate=0.02)
    rs = ReadSet.from_list([rs[i][: 1 + (i * 7) % 40] for i in range(len(rs))])
    d = N.GraphDesc.from_buffer_copy(ds.desc())
    d.match_score, d.subst_score, d.del_score, d.ins_score = costs
    ref = O.consga_scores(d, rs)
    eng = Engine(ds, desc=d)
    _say("costs", costs)
    assert np.array_equal(eng.align_combo(rs), ref)
    assert np.array_equal(eng.score_batch(rs), ref)
    assert np.array_equal(eng.score_batch(rs, slots=96, use_graphs=True), ref)
    for i in (0, 5, 17):
        assert np.array_equal(eng.last_row(rs[i]), O.consga_last_row(d, rs[i]))
    eng.close()


@pytest.mark.parametrize("sched", [(32, 1, 1), (64, 1, 2), (512, 1, 8), (128, 10 ** 9, 4), (4096, 16, 4)])
def test_batch_schedules_are_exact(sched):
    """Every tile size and sweep launch shape gives the oracle's scores."""
    ds, rs = small_workload(seed=37, n_reads=70, length=60, error=0.06, n_rate=0.02)
    ref = O.consga_scores(ds.desc(), rs)
    eng = Engine(ds)
    eng.set_schedule(*sched)
    assert eng.schedule() == dict(tile_chars=sched[0], big_level=sched[1], big_vpw=sched[2])
    assert np.array_equal(eng.score_batch(rs), ref)
    assert np.array_equal(eng.score_batch(rs, use_graphs=True), ref)
    eng.close()


def test_autotune_keeps_the_fastest_schedule():
    """The tuner keeps the fastest candidate up to timing noise: in trial
    order, a candidate displaces the kept schedule only when it is more than
    1% faster, so the kept one is within 1/0.99 of the fastest."""
    ds, rs = small_workload(seed=38, n_reads=40, length=80, error=0.04)
    eng = Engine(ds)
    tried = eng.autotune(rs, ticks=8)
    assert len(tried) >= 10 and all(t["us_per_tick"] > 0 for t in tried)
    kept = tried[0]
    for t in tried[1:]:
        if t["us_per_tick"] < 0.99 * kept["us_per_tick"]:
            kept = t
    keys = ("tile_chars", "big_level", "big_vpw")
    assert eng.schedule() == {k: kept[k] for k in keys}
    assert kept["us_per_tick"] <= min(t["us_per_tick"] for t in tried) / 0.99
    assert np.array_equal(eng.score_batch(rs), O.consga_scores(ds.desc(), rs))
    with pytest.raises(N.NativeError):
        eng.set_schedule(16, 1, 4)
    with pytest.raises(N.NativeError):
        eng.set_schedule(128, 1, 3)
    eng.close()


def test_first_chars_with_three_in_neighbours(tmp_path):
    """A string with three predecessors takes the general row kernel of the
    batched path; its scores equal the oracle's."""
    rng = np.random.default_rng(9)
    strings = [R.rbytes(rng, int(rng.integers(1, 6))) for _ in range(12)]
    edges = [(0, 3), (1, 3), (2, 3), (3, 4), (3, 5), (4, 6), (5, 6), (6, 7), (7, 8), (8, 9), (2, 9), (9, 10),
             (10, 11), (4, 11)]
    p = R.write_dataset(tmp_path, strings, edges, "tri")
    ds = Dataset.load(p["char.adj"], p["str.adj"], None, p["labels"], p["label_offsets"])
    reads = ReadSet.from_list([R.rbytes(rng, int(rng.integers(1, 30))) for _ in range(70)])
    ref = O.consga_scores(ds.desc(), reads)
    eng = Engine(ds)
    assert np.array_equal(eng.score_batch(reads), ref)
    assert np.array_equal(eng.score_batch(reads, slots=128, use_graphs=True), ref)
    assert np.array_equal(eng.align_combo(reads), ref)
    eng.close()


def test_multi_component_and_long_strings(tmp_path):
    rng = np.random.default_rng(4)
    strings, edges = [], []
    for comp in range(3):  # three components, one with a 3000-char string vertex
        base = len(strings)
        s, e = R.random_string_dag(rng, 25, 6, p_edge=0.08)
        if comp == 1:
            s[5] = R.rbytes(rng, 3000)
        strings += s
        edges += [(a + base, b + base) for a, b in e]
    p = R.write_dataset(tmp_path, strings, edges, "mc")
    ds = Dataset.load(p["char.adj"], p["str.adj"], None, p["labels"], p["label_offsets"])
    assert ds.info()["num_cc"] == 3
    reads = ReadSet.from_list([R.rbytes(rng, int(rng.integers(1, 80))) for _ in range(40)] +
                              [strings[25 + 5][100:400]])
    ref = O.consga_scores(ds.desc(), reads)
    eng = Engine(ds)
    assert np.array_equal(eng.align_combo(reads), ref)
    assert np.array_equal(eng.score_batch(reads), ref)
    for i in (0, len(reads) - 1):
        assert np.array_equal(eng.last_row(reads[i]), O.consga_last_row(ds.desc(), reads[i]))
    eng.close()


def test_golden_fixtures():
    files = sorted(GOLDEN.glob("*.json"))
    assert files, "golden fixtures missing"
    for f in files:
        g = json.loads(f.read_text())
        _say("golden", f.name)
        gp = dict(g["graph"])
        for k in ("indel_len", "sv_len"):
            if k in gp:
                gp[k] = tuple(gp[k])
        ds = Dataset.generate(**gp)
        rs = ReadSet.from_list([r.encode() for r in g["reads"]])
        eng = Engine(ds, g["alphabet"])
        assert list(eng.align_combo(rs)) == g["scores"]
        assert list(eng.score_batch(rs)) == g["scores"]
        row = eng.last_row(rs[0])
        assert int(row.sum()) == g["row0_sum"] and int(row.argmin()) == g["row0_argmin"]
        eng.close()


def test_c1_all_reads_match_oracle(c1):
    w, ds, rs = c1
    ref = O.consga_scores(ds.desc(w.alphabet), rs)
    eng = Engine(ds, w.alphabet)
    assert np.array_equal(eng.align_combo(rs), ref)
    assert np.array_equal(eng.score_batch(rs, order=plan_order(rs, ds)), ref)
    eng.close()


def test_c3_full_size_properties():
    """Full-size graph: scores agree between the two GPU paths, lie in [0, |r|],
    and error-free reads spelled along graph paths score 0."""
    w = WORKLOADS["c3"]
    ds = w.dataset()
    rs = w.readset(ds, n_reads=3)
    short = ReadSet.from_list([rs[i][:400] for i in range(3)])
    clean = ReadSet.generate(ds, seed=99, n_reads=2, length=300)
    both = ReadSet.from_list([short[i] for i in range(3)] + [clean[0], clean[1]])
    eng = Engine(ds, w.alphabet)
    _say("c3 plain")
    p = eng.align_combo(both)
    _say("c3 batched")
    b = eng.score_batch(both)
    assert np.array_equal(p, b)
    assert (p >= 0).all() and (p <= both.lengths()).all()
    assert p[3] == 0 and p[4] == 0
    eng.close()


def test_c5_large_graph_matchN(c5_ds):
    """The largest workload's graph (116.7M chars, 28 CH levels, structural
    bubbles, matchN): row offsets exceed 2^31 elements, so every index is
    64-bit.  Batched and plain scores agree on short reads; every score lies in
    [0, |r|]; error-free reads along graph paths score 0."""
    w, ds = c5_ds
    rs = w.readset(ds, n_reads=10)
    reads = ReadSet.from_list([rs[i][:1500] for i in range(10)])
    clean = ReadSet.generate(ds, seed=7, n_reads=2, length=500)
    both = ReadSet.from_list([reads[i] for i in range(10)] + [clean[0], clean[1]])
    eng = Engine(ds, w.alphabet)
    _say("c5 batched", both.lengths().tolist())
    b = eng.score_batch(both)
    assert (b >= 0).all() and (b <= both.lengths()).all()
    assert b[10] == 0 and b[11] == 0
    short = [i for i in range(len(both)) if both.lengths()[i] <= 300][:3]
    _say("c5 plain", short)
    p = eng.align_combo(both.subset(short))
    assert np.array_equal(p, b[short])
    eng.close()


def test_drop_in_drivers(tmp_path):
    """The ConSGA_* binaries on written input files: the score line equals the
    oracle's, the traceback driver's verification passes, and flags the
    reference rejects are rejected."""
    import subprocess
    from chsga_amd import WORKLOADS
    root = Path(__file__).resolve().parent.parent
    bin_ = root / "chsga_amd" / "bin"
    w = WORKLOADS["c1"]
    ds = w.dataset()
    pre = str(tmp_path / "c1")
    ds.save(pre)
    rs = w.readset(ds, n_reads=6)
    with open(pre + ".reads.fq", "wb") as f:
        for i in range(len(rs)):
            f.write(b"@r%d\n%s\n+\n%s\n" % (i, rs[i], b"I" * len(rs[i])))
    args = ["-I", pre + ".char.adj", "-i", pre + ".str.adj", "-c", pre + ".ch", "-r", pre + ".reads.fq",
            "-l", pre + ".labels", "-o", pre + ".label_offsets"]
    ref = O.consga_scores(ds.desc(), rs)
    out = subprocess.run([str(bin_ / "ConSGA_mismatchN_dag"), *args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert lines[0].startswith("input ch time: ") and lines[1].startswith("input time: ")
    assert lines[2] == "Average times (skipping first read), 95% CI:"
    assert lines[3].startswith(" Total: ") and "(n=5, df=4)" in lines[3]
    assert lines[-1] == "score: " + " ".join(str(int(x)) for x in ref)
    out = subprocess.run([str(bin_ / "ConSGA_mismatchN_dag"), *args, "-n", "5"], capture_output=True, text=True,
                         timeout=300)
    assert out.stdout.splitlines()[-1] == "score: " + " ".join(str(int(x)) for x in ref[:5])
    out = subprocess.run([str(bin_ / "ConSGA_traceback_mismatchN_dag"), *args, "-k", "100"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr
    assert "[RESULT] ALL CHECKS PASSED" in out.stdout and "(checkpoint every k=100 rows" in out.stdout
    out = subprocess.run([str(bin_ / "ConSGA_traceback_mismatchN_dag"), *args, "-k", "100", "-B"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr
    assert "[RESULT] ALL CHECKS PASSED" in out.stdout
    out = subprocess.run([str(bin_ / "ConSGA_traceback_mismatchN_dag"), *args, "-b"], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0 and " Substitutions: NA sec" in out.stdout
    # -E: ParSGA's exact insertion closure
    exact = O.exact_scores(ds.desc(), rs)
    out = subprocess.run([str(bin_ / "ConSGA_mismatchN_dag"), *args, "-E", "-B"], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.splitlines()[-1] == "score: " + " ".join(str(int(x)) for x in exact)
    out = subprocess.run([str(bin_ / "ConSGA_traceback_mismatchN_dag"), *args, "-k", "100", "-B", "-E"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr
    assert "[RESULT] ALL CHECKS PASSED" in out.stdout
    # -X: the same scores, rows closed on match-cut components (plain path)
    out = subprocess.run([str(bin_ / "ConSGA_mismatchN_dag"), *args, "-X"], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.splitlines()[-1] == "score: " + " ".join(str(int(x)) for x in exact)
    out = subprocess.run([str(bin_ / "ConSGA_traceback_mismatchN_dag"), *args, "-k", "100", "-X"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr
    assert "[RESULT] ALL CHECKS PASSED" in out.stdout
    bad = subprocess.run([str(bin_ / "ConSGA_mismatchN_dag"), *args, "-a", "x"], capture_output=True, text=True)
    assert bad.returncode != 0 and "Usage" in bad.stderr


def test_int16_storage_equals_int32_storage():
    """Scores stored as int16 (used when every stored value provably stays
    below 2^15) equal int32 storage and the oracle; a read longer than 32767
    chars takes the int32 path."""
    from chsga_amd import native as N
    ds, rs = small_workload(seed=14, n_reads=40, length=200, n_rate=0.02)
    ref = O.consga_scores(ds.desc(), rs)
    eng = Engine(ds)
    for narrow in (0, 1):
        eng.set_option(N.OPT_NARROW, narrow)
        assert np.array_equal(eng.align_combo(rs), ref)
        assert np.array_equal(eng.score_batch(rs), ref)
    long_read = ReadSet.generate(ds, seed=5, n_reads=1, length=33000, sub=0.05, ins=0.02, dele=0.02)
    assert len(long_read[0]) == 33000
    lref = O.consga_scores(ds.desc(), long_read)
    assert lref[0] > 0
    eng.set_option(N.OPT_NARROW, 1)
    assert np.array_equal(eng.align_combo(long_read), lref)
    eng.close()


def test_top_levels_as_one_closure_equal_level_by_level():
    """The batched sweep's top CH levels run as one launch over precomputed
    {source, summed weight} lists (the default, also in the plain path's
    full sweeps); scores and traceback paths equal the level-by-level
    sweep's and the oracle's: on c1 (every level
    above 0 in the closure), on small graphs of both alphabets, packed and
    char rows, and with other insertion costs (the weights' multiplier)."""
    w = WORKLOADS["c1"]
    ds2, rs2 = small_workload(seed=53, n_reads=80, length=150, error=0.08, n_rate=0.03)
    cases = [(w.dataset(), w.alphabet, None, None), (ds2, N.ALPHABET_MISMATCH_N, rs2, None),
             (ds2, N.ALPHABET_MATCH_N, rs2, None)]
    for ins in (0, 2):
        d = ds2.desc()
        d.ins_score = ins
        cases.append((ds2, N.ALPHABET_MISMATCH_N, rs2, d))
    for ds, alpha, rs, desc in cases:
        if rs is None:
            rs = w.readset(ds, n_reads=40)
        rs = ReadSet.from_list([rs[i][: 1 + (i * 31) % len(rs[i])] for i in range(len(rs))])
        for packed in (True, False):
            a = Engine(ds, alpha, desc=desc, packed=packed)
            _say("top closure vs level by level", alpha, packed, None if desc is None else desc.ins_score)
            ref = O.consga_scores(desc if desc is not None else ds.desc(alpha), rs)
            sa, ta = a.score_batch(rs, use_graphs=True), a.traceback_batch(rs, k=40)
            pa = a.align_combo(rs)  # the plain path's full sweeps take the same lists
            a.set_option(N.OPT_TOP_CLOSURE, 0)
            sb, tb = a.score_batch(rs, use_graphs=True), a.traceback_batch(rs, k=40)
            pb = a.align_combo(rs)
            assert np.array_equal(sa, ref) and np.array_equal(sb, ref)
            assert np.array_equal(pa, ref) and np.array_equal(pb, ref)
            for i in range(len(rs)):
                assert ta[i].score == tb[i].score and np.array_equal(ta[i].ops, tb[i].ops), i
            a.close()


def test_packed_rows_equal_char_rows():
    """Batched rows packed as 2-bit deltas (the default on compact graphs
    with unit costs) give the scores and paths of one score per char, on c1,
    on small graphs with N and on reads of uneven length; a graph with other
    costs keeps char rows."""
    from chsga_amd import WORKLOADS
    w = WORKLOADS["c1"]
    cases = [(w.dataset(), w.alphabet, None)]
    ds2, rs2 = small_workload(seed=51, n_reads=90, length=150, error=0.08, n_rate=0.03)
    cases.append((ds2, N.ALPHABET_MISMATCH_N, rs2))
    cases.append((ds2, N.ALPHABET_MATCH_N, rs2))
    for ds, alpha, rs in cases:
        if rs is None:
            rs = w.readset(ds, n_reads=40)
        rs = ReadSet.from_list([rs[i][: 1 + (i * 29) % len(rs[i])] for i in range(len(rs))])
        a = Engine(ds, alpha)
        b = Engine(ds, alpha, packed=False)
        a.set_option(N.OPT_PACKED_ROWS, 1)
        b.set_option(N.OPT_PACKED_ROWS, 0)
        _say("packed vs char rows", alpha, len(rs))
        sa, sb = a.score_batch(rs), b.score_batch(rs)
        assert np.array_equal(sa, sb)
        assert np.array_equal(sa, a.align_combo(rs))
        ta, tb = a.traceback_batch(rs, k=40), b.traceback_batch(rs, k=40)
        for i in range(len(rs)):
            assert ta[i].score == tb[i].score and np.array_equal(ta[i].ops, tb[i].ops), i
        a.close()
        b.close()
    d = ds2.desc()
    d.del_score = 2
    e = Engine(ds2, desc=d)
    with pytest.raises(N.NativeError):
        e.set_option(N.OPT_PACKED_ROWS, 1)
    e.close()
