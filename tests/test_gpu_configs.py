"""GPU parity at the sizes of the five workloads (c1 is covered read by read in
test_gpu_parity.py).

The oracle (O2, the C++/OpenMP restatement of src/align.hpp:470-531) runs on
the host in a background thread while the GPU works, on read prefixes it
finishes in seconds at these graph sizes; full-length reads are checked by
the size-independent properties (plain == batched, 0 <= score <= |r|, every
traceback path valid and rescored to its score, the verification of
src/ConSGA_traceback.cpp:184-225).
"""
from concurrent.futures import ThreadPoolExecutor
import time

import numpy as np
import pytest

from chsga_amd import Engine, ReadSet, WORKLOADS
from chsga_amd import native as N
from chsga_amd.api import verify_path
from chsga_amd.scheduler import plan_order, simulate_ticks
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _cut(rs, n):
    return ReadSet.from_list([rs[i][:n] for i in range(len(rs))])


@pytest.fixture(scope="module")
def pool():
    with ThreadPoolExecutor(max_workers=1) as ex:
        yield ex


def _check_paths(desc, reads, paths, scores):
    for i, p in enumerate(paths):
        assert p.score == scores[i], i
        rc, bad = O.validate_path(desc, reads[i], p.ops, p.score)
        assert rc == 0, (i, rc, bad)
        assert verify_path(desc, reads[i], p.ops, p.score) == (0, -1), i


@pytest.mark.timeout(600)
def test_c2_plain_path_against_oracle(pool):
    """c2: 5 Mbp DAG (~5.2M chars), 10 kbp reads at 5% error (2:1:1).  The
    plain path (one read per launch sequence) equals O2 on 2 kbp prefixes of
    4 reads; at full length plain == batched and 0 <= s <= |r|."""
    w = WORKLOADS["c2"]
    ds = w.dataset()
    rs = w.readset(ds, n_reads=4)
    assert (rs.lengths() == 10_000).all()
    d = ds.desc(w.alphabet)
    pre = _cut(rs, 2000)
    fut = pool.submit(O.consga_scores, d, pre)
    eng = Engine(ds, w.alphabet)
    _say("c2 plain, 2 kbp prefixes")
    p_pre = eng.align_combo(pre)
    b_pre = eng.score_batch(pre)
    _say("c2 plain, full length")
    t = time.perf_counter()
    p = eng.align_combo(rs)
    _say(f"c2 plain: {rs.lengths().sum() * d.V / (time.perf_counter() - t) / 1e9:.1f} GCUPS")
    b = eng.score_batch(rs, use_graphs=True)
    assert np.array_equal(p, b)
    assert (p >= 0).all() and (p <= rs.lengths()).all() and (p > 0).all()
    ref = fut.result()
    _say("c2 oracle", ref.tolist(), "gpu", p_pre.tolist())
    assert np.array_equal(p_pre, ref) and np.array_equal(b_pre, ref)
    eng.close()


@pytest.mark.timeout(600)
def test_c4_slice_learned_and_formula_orders():
    """c4: the c2 graph, a 1,024-read slice of the 16,384-read stream
    (log-normal 0.5-30 kbp).  Queue orders from the learned cost model and
    from the plain formula give identical scores, equal to the plain path on a
    subset; the engine's tick count equals the queue simulation of each
    order (the device queue runs exactly the planned schedule)."""
    from chsga_amd import learned as L
    w = WORKLOADS["c4"]
    ds = w.dataset()
    rs = w.readset(ds, n_reads=1024)
    model = L.load_default()
    assert model is not None, "the trained cost model is committed under chsga_amd/data"
    eng = Engine(ds, w.alphabet)
    eng.set_option(N.OPT_FORWARD_TAIL, 0)  # every tick batched: the tick count is the queue's
    res = {}
    for name, m in (("learned", model), ("formula", None)):
        order = plan_order(rs, ds, m)
        t = time.perf_counter()
        sc = eng.score_batch(rs, order=order, use_graphs=True)
        dt = time.perf_counter() - t
        ticks = eng.last_stats["rows"]
        res[name] = (sc, ticks, dt, order)
        _say(f"c4 {name} order: {ticks} ticks (simulated {simulate_ticks(rs.lengths(), order)}), {dt:.2f} s, "
             f"{rs.lengths().sum() * ds.info()['V'] / dt / 1e9:.1f} GCUPS")
        assert ticks == simulate_ticks(rs.lengths(), order)
    assert np.array_equal(res["learned"][0], res["formula"][0])
    sc = res["learned"][0]
    eng.set_option(N.OPT_FORWARD_TAIL, 3)
    _say("c4 learned order with the forward tail on the plain path")
    assert np.array_equal(eng.score_batch(rs, order=res["learned"][3], use_graphs=True), sc)
    lens = rs.lengths()
    assert (sc >= 0).all() and (sc <= lens).all()
    sub = sorted(set(np.argsort(lens, kind="stable")[:3].tolist() + [100, 700]))
    _say("c4 plain subset", lens[sub].tolist())
    assert np.array_equal(eng.align_combo(rs.subset(sub)), sc[sub])
    # the learned order packs no worse than the formula's
    assert res["learned"][1] <= res["formula"][1]
    eng.close()
    model.close()


@pytest.fixture(scope="module")
def c3():
    w = WORKLOADS["c3"]
    ds = w.dataset()
    return w, ds, w.readset(ds, n_reads=128)


@pytest.mark.timeout(600)
def test_c3_prefixes_against_oracle_and_batched_traceback(c3, pool):
    """c3 graph (20 Mbp, ~22M chars, mismatchN, 1% N): 64 reads of the config
    distribution cut to 400 chars.  Scores of the plain path, the batched
    path and the batched traceback equal O2 on 8 of them (all 64 agree among
    the GPU paths); the batched traceback at k = 50 (packed rows, the
    capture path for 16 or more lanes, the LDS-transposed encode) gives valid
    paths equal to the per-read traceback's."""
    w, ds, rs = c3
    d = ds.desc(w.alphabet)
    reads = _cut(rs.subset(range(64)), 400)
    fut = pool.submit(O.consga_scores, d, reads.subset(range(8)))
    eng = Engine(ds, w.alphabet)
    assert eng.packed
    _say("c3 400-char prefixes: plain")
    p = eng.align_combo(reads)
    _say("c3 batched")
    b = eng.score_batch(reads, order=plan_order(reads, ds), use_graphs=True)
    assert np.array_equal(p, b)
    _say("c3 batched traceback k=50")
    out = eng.traceback_batch(reads, k=50, use_graphs=True)
    _check_paths(d, reads, out, b)
    for i in range(0, 64, 8):
        _say("c3 per-read traceback", i)
        tb = eng.align_read_with_traceback(reads[i], k=50)
        assert tb.score == out[i].score and np.array_equal(tb.ops, out[i].ops), i
    ref = fut.result()
    _say("c3 oracle", ref.tolist(), "gpu", p[:8].tolist())
    assert np.array_equal(p[:8], ref)
    eng.close()


@pytest.mark.timeout(600)
def test_c3_pasgal_against_oracle(c3, pool):
    """PaSGAL_match (src/PaSGALoverrides/align.hpp) on the c3 graph: two reads
    cut to 80 chars, the second reverse-complemented.  Score, strand, ends,
    starts and CIGAR of each equal O5 (phase 1 on both strands, phase 1-R,
    phase 2), which runs on the host meanwhile; a read whose pinned end cell
    the closure raised (CHSGA_PASGAL_PIN_MOVED, possible with this workload's
    N chars) is compared on its phase-1 results."""
    from chsga_amd.api import PasgalAligner
    comp = bytes.maketrans(b"ACGTN", b"TGCAN")
    w, ds, rs = c3
    reads = ReadSet.from_list([rs[0][:80], rs[1][:80].translate(comp)[::-1]])
    fut = pool.submit(O.pasgal_align, ds.desc(w.alphabet), reads)
    al = PasgalAligner(ds)
    got = al.align(reads)
    al.close()
    exp = fut.result()
    _say("c3 PaSGAL", [(g["score"], g["strand"], g["qry_start"], g["qry_end"], g["flags"]) for g in got])
    phase1 = ("score", "strand", "qry_end", "ref_end", "score_fwd", "score_rc", "qry_end_fwd", "ref_end_fwd",
              "qry_end_rc", "ref_end_rc")
    for i, (g, e) in enumerate(zip(got, exp)):
        if g["flags"] & N.PASGAL_PIN_MOVED:
            assert b"N" in reads[i], i
            assert all(g[k] == e[k] for k in phase1), (i, g, e)
        else:
            assert g == e, (i, g, e)
    assert got[1]["strand"] == "-"


@pytest.mark.timeout(600)
def test_c3_full_length_batched_traceback_validates(c3):
    """The headline work at full length: 64 c3 reads (log-normal, median
    3 kbp), batched traceback with checkpoint rows every 500 in HBM; every
    path is valid and rescores to the read's batched score, which equals the
    plain path's on the two shortest reads."""
    w, ds, rs = c3
    d = ds.desc(w.alphabet)
    reads = rs.subset(range(64, 128))
    eng = Engine(ds, w.alphabet)
    _say("c3 full-length scores", int(reads.lengths().sum()), int(reads.lengths().max()))
    sc = eng.score_batch(reads, order=plan_order(reads, ds), use_graphs=True)
    _say("c3 full-length batched traceback k=500")
    t = time.perf_counter()
    out = eng.traceback_batch(reads, k=500, order=plan_order(reads, ds), use_graphs=True)
    _say(f"c3 traceback: {time.perf_counter() - t:.1f} s, {eng.last_stats['rows']} ticks, "
         f"{eng.last_stats['walk_fallbacks']} staged walks")
    _check_paths(d, reads, out, sc)
    short = np.argsort(reads.lengths(), kind="stable")[:2].tolist()
    assert np.array_equal(eng.align_combo(reads.subset(short)), sc[short])
    eng.close()


@pytest.mark.timeout(600)
def test_c5_prefixes_against_oracle(c5_ds, pool):
    """c5 graph (100 Mbp, ~117M chars, structural bubbles, matchN): two reads
    of the mixture cut to 200 chars; plain and batched scores equal O2."""
    w, ds = c5_ds
    rs = w.readset(ds, n_reads=6)
    longs = [i for i in range(len(rs)) if rs.lengths()[i] >= 200][:2]
    reads = _cut(rs.subset(longs), 200)
    d = ds.desc(w.alphabet)
    fut = pool.submit(O.consga_scores, d, reads)
    eng = Engine(ds, w.alphabet)
    _say("c5 plain", reads.lengths().tolist())
    p = eng.align_combo(reads)
    b = eng.score_batch(reads)
    assert np.array_equal(p, b)
    ref = fut.result()
    _say("c5 oracle", ref.tolist(), "gpu", p.tolist())
    assert np.array_equal(p, ref)
    eng.close()


@pytest.mark.timeout(600)
def test_c5_long_read(c5_ds):
    """c5 at the workload's read lengths: the stream's first long read (of
    the 5-50 kbp half, up to 5% error) cut to 16 kbp, on the 116.7M char
    graph.  Its score is at most |r|; the per-read traceback (k = 500) with
    its checkpoint rows in pinned host memory and the two-level segment
    store, as the workload's traceback store spills, gives a path that
    validates and rescores to that score; an error-free 6 kbp read sampled
    from the graph scores 0.  (The whole 34,376-char read passed the same
    checks in 161 s; the cut keeps the GPU suite short.)"""
    w, ds = c5_ds
    rs = w.readset(ds, n_reads=8)
    i = int(np.argmax(rs.lengths()))
    assert rs.lengths()[i] >= 16000
    clean = ReadSet.generate(ds, seed=11, n_reads=1, length=6000)
    reads = ReadSet.from_list([rs[i][:16000], clean[0]])
    d = ds.desc(w.alphabet)
    eng = Engine(ds, w.alphabet)
    _say("c5 plain", reads.lengths().tolist())
    s = eng.align_combo(reads)
    assert 0 <= s[0] <= len(reads[0]) and s[1] == 0
    eng.set_option(N.OPT_CKPT_HBM_ROWS, 0)
    eng.set_option(N.OPT_SEG_ROWS, 64)
    _say("c5 scores", s.tolist(), "traceback k=500, host checkpoints, two-level segments")
    tb = eng.align_read_with_traceback(reads[0], k=500)
    assert eng.last_stats["host_bytes"] > 0
    _check_paths(d, reads.subset([0]), [tb], s[:1])
    eng.close()


@pytest.mark.timeout(600)
def test_gfa_chain_through_the_drivers_against_oracle(tmp_path):
    """A graph entering as GFA goes through the reference's tool chain:
    gfa_to_csr (src/handlegraph_to_csr.cpp's two CSR texts), csr_to_adj on
    both (src/csr_to_adj.cpp:39-114), build_pch -v, then the drop-in drivers
    on those files.  The score line equals O2 on the dataset the same files
    load as, and the traceback driver's verification passes."""
    import subprocess
    from pathlib import Path
    from chsga_amd import Dataset
    bin_ = Path(__file__).resolve().parent.parent / "chsga_amd" / "bin"
    src = Dataset.generate(seed=17, backbone_bp=40_000, snp_rate=1 / 30, indel_rate=1 / 200, indel_len=(1, 15))
    a = src.arrays()
    lab = src.labels()
    n = a["n_s"]
    lines = [f"S\t{i + 1}\t{lab[a['str_off'][i]:a['str_off'][i + 1]].decode()}" for i in range(n)]
    for i in range(n):
        for j in a["str_out_dst"][a["str_out_off"][i]:a["str_out_off"][i + 1]]:
            lines.append(f"L\t{i + 1}\t+\t{int(j) + 1}\t+\t0M")
    g = tmp_path / "g.gfa"
    g.write_text("\n".join(lines) + "\n")
    pre = str(tmp_path / "g")
    run = lambda *c: subprocess.run([str(x) for x in c], check=True, capture_output=True, text=True, timeout=300)
    run(bin_ / "gfa_to_csr", g, pre + ".str.csr", pre + ".char.csr")
    run(bin_ / "csr_to_adj", pre + ".str.csr", pre + ".str.adj", pre + ".labels", pre + ".label_offsets")
    run(bin_ / "csr_to_adj", pre + ".char.csr", pre + ".char.adj", pre + ".char.labels", pre + ".char.label_offsets")
    run(bin_ / "build_pch", "-i", pre + ".str.adj", "-o", pre + ".ch", "-v", pre + ".label_offsets")
    rs = ReadSet.generate(src, seed=5, n_reads=12, length=300, sub=0.02, ins=0.01, dele=0.01, n_rate=0.01)
    with open(pre + ".reads.fq", "wb") as f:
        for i in range(len(rs)):
            f.write(b"@r%d\n%s\n+\n%s\n" % (i, rs[i], b"I" * len(rs[i])))
    ds = Dataset.load(pre + ".char.adj", pre + ".str.adj", pre + ".ch", pre + ".labels", pre + ".label_offsets")
    assert ds.info()["n_s"] == 2 * n
    ref = O.consga_scores(ds.desc(), rs)
    args = ["-I", pre + ".char.adj", "-i", pre + ".str.adj", "-c", pre + ".ch", "-r", pre + ".reads.fq",
            "-l", pre + ".labels", "-o", pre + ".label_offsets"]
    for name in ("ConSGA_mismatchN_dag", "ConSGA_mismatchN"):
        _say("driver", name)
        out = run(bin_ / name, *args)
        assert out.stdout.splitlines()[-1] == "score: " + " ".join(str(int(x)) for x in ref)
    _say("traceback driver")
    out = run(bin_ / "ConSGA_traceback_mismatchN_dag_800gb", *args, "-k", "50", "-B")
    assert "[RESULT] ALL CHECKS PASSED" in out.stdout
    assert "score=" + str(int(ref[0])) in out.stdout
