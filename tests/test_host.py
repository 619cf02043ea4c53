"""Host side without a GPU: generator, CH builder, formats, read loader, C ABI."""
import gzip
import re
from pathlib import Path

import numpy as np
import pytest

from chsga_amd import Dataset, ReadSet, WORKLOADS
from chsga_amd import native as N
from tests import refmodels as R

ROOT = Path(__file__).resolve().parent.parent


def test_generator_is_deterministic():
    a = Dataset.generate(seed=4, backbone_bp=5000, snp_rate=1 / 30, indel_rate=1 / 200, indel_len=(1, 9))
    b = Dataset.generate(seed=4, backbone_bp=5000, snp_rate=1 / 30, indel_rate=1 / 200, indel_len=(1, 9))
    c = Dataset.generate(seed=5, backbone_bp=5000, snp_rate=1 / 30, indel_rate=1 / 200, indel_len=(1, 9))
    assert a.labels() == b.labels() and a.labels() != c.labels()
    ra = ReadSet.generate(a, seed=9, n_reads=5, length=200, sub=0.02, ins=0.01, dele=0.01, n_rate=0.01)
    rb = ReadSet.generate(b, seed=9, n_reads=5, length=200, sub=0.02, ins=0.01, dele=0.01, n_rate=0.01)
    assert ra.data == rb.data and (ra.lengths() == 200).all()


def test_c1_shape():
    """c1: 100 kbp backbone, SNPs 1/50, indels 1/500 -> ~1.03-1.1e5 chars."""
    w = WORKLOADS["c1"]
    info = w.dataset().info()
    assert 100_000 <= info["V"] <= 110_000
    assert info["num_cc"] == 1


def test_lognormal_read_lengths():
    ds = Dataset.generate(seed=1, backbone_bp=200_000, snp_rate=1 / 25, indel_rate=1 / 200, indel_len=(1, 50))
    rs = ReadSet.generate(ds, seed=3, n_reads=400, length=3000, len_mode=1, len_sigma=0.6, len_min=1000,
                          len_max=30000, sub=0.01, ins=0.005, dele=0.005, n_rate=0.01)
    L = rs.lengths()
    assert L.min() >= 1000 and L.max() <= 30000
    assert 2500 < np.median(L) < 3600
    assert 0.003 < rs.data.count(b"N") / len(rs.data) < 0.02


def _ch_distance(a, s, t, lens):
    """Up-down distance in the CH between strings s and t (the PHAST query the
    sweep evaluates), with vertex weights len-1 on intermediate vertices."""
    rank, up_off, up_v, up_w = a["rank"], a["up_off"], a["up_v"], a["up_w"]
    dn_off, dn_v, dn_w = a["dn_off"], a["dn_v"], a["dn_w"]
    n = len(rank)
    og = np.empty(n, np.int64)
    og[rank] = np.arange(n)
    INF = 1 << 40
    # forward search over up-edges from s
    fwd = {rank[s]: 0}
    stack = [rank[s]]
    while stack:
        x = stack.pop()
        for e in range(up_off[x], up_off[x + 1]):
            y = up_v[e]
            d = fwd[x] + (lens[og[x]] - 1 if x != rank[s] else 0) + up_w[e]
            if d < fwd.get(y, INF):
                fwd[y] = d
                stack.append(y)
    # backward search from t over down-edges reversed (y -> x stored at x)
    bwd = {rank[t]: 0}
    stack = [rank[t]]
    while stack:
        x = stack.pop()
        for e in range(dn_off[x], dn_off[x + 1]):
            y = dn_v[e]
            d = bwd[x] + (lens[og[x]] - 1 if x != rank[t] else 0) + dn_w[e]
            if d < bwd.get(y, INF):
                bwd[y] = d
                stack.append(y)
    best = INF
    for v, d in fwd.items():
        if v in bwd:
            mid = lens[og[v]] - 1 if (v != rank[s] and v != rank[t]) else 0
            best = min(best, d + bwd[v] + mid)
    return best


def _true_distance(n, out, lens, s, t):
    import heapq
    INF = 1 << 40
    dist = [INF] * n
    dist[s] = 0
    pq = [(0, s)]
    while pq:
        d, x = heapq.heappop(pq)
        if d > dist[x]:
            continue
        for y in out[x]:
            nd = d + (lens[x] - 1 if x != s else 0) + 1
            if nd < dist[y]:
                dist[y] = nd
                heapq.heappush(pq, (nd, y))
    return dist[t]


def test_ch_is_a_valid_vertex_weighted_hierarchy(tmp_path):
    """Edges go to strictly higher layers, every shortcut decomposes into its two
    recorded sub-edges with w = w1 + w2 + len(mid) - 1, and the up-down CH
    distance equals the true vertex-weighted shortest path distance."""
    rng = np.random.default_rng(2)
    for t in range(8):
        strings, edges = R.random_string_dag(rng, 40, 4, p_edge=0.06)
        p = R.write_dataset(tmp_path, strings, edges, f"h{t}")
        ds = Dataset.load(p["char.adj"], p["str.adj"], None, p["labels"], p["label_offsets"])
        a = ds.arrays()
        n = len(strings)
        lens = [len(s) for s in strings]
        layer = np.zeros(n, np.int64)
        for l in range(len(a["layer_off"]) - 1):
            layer[a["layer_off"][l]:a["layer_off"][l + 1]] = l
        og = np.empty(n, np.int64)
        og[a["rank"]] = np.arange(n)
        for x in range(n):
            for e in range(a["up_off"][x], a["up_off"][x + 1]):
                y = a["up_v"][e]
                assert layer[y] > layer[x]
                m = a["up_mid"][e]
                if m >= 0:  # shortcut x -> y through m: down-edge (x -> m) at m and up-edge (m -> y) at m
                    w1 = [a["dn_w"][k] for k in range(a["dn_off"][m], a["dn_off"][m + 1]) if a["dn_v"][k] == x]
                    w2 = [a["up_w"][k] for k in range(a["up_off"][m], a["up_off"][m + 1]) if a["up_v"][k] == y]
                    assert w1 and w2 and a["up_w"][e] == w1[0] + w2[0] + lens[og[m]] - 1
            for e in range(a["dn_off"][x], a["dn_off"][x + 1]):
                assert layer[a["dn_v"][e]] > layer[x]
        out = [[] for _ in range(n)]
        for u, v in edges:
            out[u].append(v)
        for _ in range(30):
            s, tt = sorted(rng.integers(0, n, size=2))
            if s == tt:
                continue
            want = _true_distance(n, out, lens, int(s), int(tt))
            if want >= (1 << 40):
                continue
            assert _ch_distance(a, int(s), int(tt), lens) == want


def test_dataset_round_trip(tmp_path):
    ds = Dataset.generate(seed=8, backbone_bp=4000, snp_rate=1 / 30, indel_rate=1 / 100, indel_len=(1, 20))
    pre = str(tmp_path / "rt")
    ds.save(pre)
    head = open(pre + ".char.adj").read().split("\n")[:3]
    assert head[0] == "AdjacencyGraph"
    ds2 = Dataset.load(pre + ".char.adj", pre + ".str.adj", pre + ".ch", pre + ".labels", pre + ".label_offsets")
    a, b = ds.arrays(), ds2.arrays()
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k
    # labels file: '\n' bytes are dropped (src/ConSGA.cpp:166-169)
    lab = open(pre + ".labels", "rb").read()
    open(pre + ".labels", "wb").write(lab[:10] + b"\n" + lab[10:] + b"\n")
    ds3 = Dataset.load(pre + ".char.adj", pre + ".str.adj", pre + ".ch", pre + ".labels", pre + ".label_offsets")
    assert ds3.labels() == ds.labels()


def test_inconsistent_inputs_are_rejected(tmp_path):
    ds = Dataset.generate(seed=8, backbone_bp=3000, snp_rate=1 / 30, indel_rate=1 / 100, indel_len=(1, 20))
    pre = str(tmp_path / "bad")
    ds.save(pre)
    lab = open(pre + ".labels", "rb").read()
    open(pre + ".labels", "wb").write(lab[:-5])
    with pytest.raises(N.NativeError, match="labels"):
        Dataset.load(pre + ".char.adj", pre + ".str.adj", pre + ".ch", pre + ".labels", pre + ".label_offsets")
    open(pre + ".labels", "wb").write(lab)
    open(pre + ".ch", "wb").write(b"not a CH file")
    with pytest.raises(N.NativeError, match="CHSGACH1"):
        Dataset.load(pre + ".char.adj", pre + ".str.adj", pre + ".ch", pre + ".labels", pre + ".label_offsets")


def test_read_loader_formats(tmp_path):
    seqs = [b"ACGTacgtNN", b"GGGTTT", b"A"]
    fa = tmp_path / "r.fa"
    fa.write_bytes(b">r1 x\nACGTac\ngtNN\n>r2\nGGGTTT\n>r3\nA\n")
    fq = tmp_path / "r.fq"
    fq.write_bytes(b"@r1\nACGTacgtNN\n+\nIIIIIIIIII\n@r2\nGGGTTT\n+r2\n######\n@r3\nA\n+\n@\n")
    gz = tmp_path / "r.fq.gz"
    gz.write_bytes(gzip.compress(fq.read_bytes()))
    txt = tmp_path / "r.txt"
    txt.write_bytes(b"ACGTacgtNN\nGGGTTT\n\nA\n")
    up = [s.upper() for s in seqs]
    for f in (fa, fq, gz, txt):
        rs = ReadSet.load(str(f))
        assert [rs[i] for i in range(len(rs))] == up, f
    assert len(ReadSet.load(str(fq), cap=2)) == 2  # the reference stops after 5 (src/load_reads.hpp:5)
    empty = tmp_path / "e.fq"
    empty.write_bytes(b"")
    with pytest.raises(N.NativeError, match="No reads"):
        ReadSet.load(str(empty))


def _declared(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(chsga_[a-z_0-9]+)\s*\(", text)))


def test_c_abi_exports_every_declared_symbol():
    import ctypes
    eng = ctypes.CDLL(str(ROOT / "chsga_amd" / "lib" / "libchsga.so"))
    host = ctypes.CDLL(str(ROOT / "chsga_amd" / "lib" / "libchsga_host.so"))
    abi = _declared(ROOT / "include" / "chsga" / "abi.h")
    hst = _declared(ROOT / "include" / "chsga" / "host.h")
    assert "chsga_score_reads" in abi and "chsga_traceback_read" in abi and "chsga_ctx_create" in abi
    for s in abi:
        assert hasattr(eng, s), s
    for s in hst:
        assert hasattr(host, s), s


def test_engine_rejects_bad_descriptor_without_touching_the_gpu():
    import ctypes
    from chsga_amd.native import engine_lib, GraphDesc
    lib = engine_lib()
    h = ctypes.c_void_p()
    d = GraphDesc()
    assert lib.chsga_ctx_create(ctypes.byref(d), 0, ctypes.byref(h)) == -1
    assert b"descriptor" in lib.chsga_last_error()
    ds = Dataset.generate(seed=8, backbone_bp=3000, snp_rate=1 / 30, indel_rate=1 / 100, indel_len=(1, 20))
    d = GraphDesc.from_buffer_copy(ds.desc())
    d.alphabet = 3
    assert lib.chsga_ctx_create(ctypes.byref(d), 0, ctypes.byref(h)) == -1
    d.alphabet = 5
    d.ins_score = -1
    assert lib.chsga_ctx_create(ctypes.byref(d), 0, ctypes.byref(h)) == -1


def test_drivers_under_every_reference_name():
    """The ConSGA targets of the reference's Makefile (:121-124, :142-145,
    :156-171) and PaSGAL_match (:197-213) are built; each rejects an unknown flag with its usage line and
    a missing input with an error, before any GPU call."""
    import subprocess
    bin_ = ROOT / "chsga_amd" / "bin"
    names = [f"ConSGA_{a}{s}" for a in ("matchN", "mismatchN") for s in ("", "_dag")] + \
        [f"ConSGA_traceback_{a}{s}" for a in ("matchN", "mismatchN") for s in ("", "_dag")] + \
        ["ConSGA_traceback_mismatchN_dag_800gb", "PaSGAL_match"]
    for n in names:
        assert (bin_ / n).exists(), n
        bad = subprocess.run([str(bin_ / n), "-a", "x"], capture_output=True, text=True, timeout=60)
        assert bad.returncode != 0 and "Usage" in bad.stderr, n
        miss = subprocess.run([str(bin_ / n), "-I", "x"], capture_output=True, text=True, timeout=60)
        assert miss.returncode != 0 and "missing required arguments" in miss.stderr, n
