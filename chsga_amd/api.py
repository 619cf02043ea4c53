"""Python host mirror of ConSGA's entry points over the C ABI.

  Dataset.generate / Dataset.load   the drivers' input stage (src/ConSGA.cpp:154-211)
  Engine.align_combo                align_combo (src/align.hpp:470-531): exact int32 scores
  Engine.align_read_with_traceback  align_read_with_traceback (src/traceback.hpp:588-1031)
  Engine.score_batch                batched align_combo over a read stream
  PasgalAligner.align               PaSGAL_match's alignToDAGLocal (src/PaSGALoverrides/align.hpp:500-700)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native as N


class Dataset:
    """String graph + char graph + contraction hierarchy held by libchsga_host."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        self._descs: dict[int, N.GraphDesc] = {}

    @classmethod
    def generate(cls, seed: int, backbone_bp: int, snp_rate: float, indel_rate: float, indel_len=(1, 10),
                 sv_rate: float = 0.0, sv_len=(50, 5000)) -> "Dataset":
        p = N.GraphParams(seed, backbone_bp, snp_rate, indel_rate, indel_len[0], indel_len[1], sv_rate,
                          sv_len[0], sv_len[1])
        h = C.c_void_p()
        N._check_host(N.host_lib().chsga_host_generate(C.byref(p), C.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, char_adj: str, str_adj: str, ch: str | None, labels: str, offsets: str) -> "Dataset":
        h = C.c_void_p()
        enc = lambda s: None if s is None else str(s).encode()
        N._check_host(N.host_lib().chsga_host_load(enc(char_adj), enc(str_adj), enc(ch), enc(labels), enc(offsets),
                                                   C.byref(h)))
        return cls(h.value)

    @classmethod
    def load_gfa(cls, path: str) -> "Dataset":
        """GFA 1 graph, both orientations of every segment
        (chsga_host_load_gfa)."""
        h = C.c_void_p()
        N._check_host(N.host_lib().chsga_host_load_gfa(str(path).encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def load_string_csr(cls, path: str) -> "Dataset":
        """The string-labelled CSR text of handlegraph_to_csr
        (chsga_host_load_string_csr)."""
        h = C.c_void_p()
        N._check_host(N.host_lib().chsga_host_load_string_csr(str(path).encode(), C.byref(h)))
        return cls(h.value)

    def reverse(self, seed: int = 1) -> "Dataset":
        """The reversed graph (char j becomes V - 1 - j) with its own CH, for
        PaSGAL's phase 1-R (chsga_host_reverse)."""
        h = C.c_void_p()
        N._check_host(N.host_lib().chsga_host_reverse(self._h, seed, C.byref(h)))
        return Dataset(h.value)

    def save(self, prefix: str) -> None:
        N._check_host(N.host_lib().chsga_host_save(self._h, str(prefix).encode()))

    def write_csr(self, string_csr: str, char_csr: str) -> None:
        N._check_host(N.host_lib().chsga_host_write_csr(self._h, str(string_csr).encode(), str(char_csr).encode()))

    def rebuild_ch(self, seed: int) -> None:
        N._check_host(N.host_lib().chsga_host_build_ch(self._h, seed))
        self._descs.clear()

    def info(self) -> dict:
        i = N.DatasetInfo()
        N._check_host(N.host_lib().chsga_host_info(self._h, C.byref(i)))
        return i.as_dict()

    def desc(self, alphabet: int = N.ALPHABET_MISMATCH_N) -> N.GraphDesc:
        if alphabet not in self._descs:
            d = N.GraphDesc()
            N._check_host(N.host_lib().chsga_host_describe(self._h, alphabet, C.byref(d)))
            self._descs[alphabet] = d
        return self._descs[alphabet]

    # numpy views (valid while the dataset lives)
    def labels(self) -> bytes:
        d = self.desc()
        return C.string_at(d.labels, d.V)

    def arrays(self) -> dict:
        d = self.desc()

        def arr(ptr, n, dt):
            if not ptr or n == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ptr, shape=(n,)).view(dt).copy()

        return {
            "V": d.V, "n_s": d.n_s,
            "labels": np.frombuffer(self.labels(), np.uint8).copy(),
            "in_off": arr(d.in_off, d.V + 1, np.uint64), "in_src": arr(d.in_src, d.E, np.uint32),
            "str_off": arr(d.str_off, d.n_s + 1, np.uint64),
            "str_out_off": arr(d.str_out_off, d.n_s + 1, np.uint64), "str_out_dst": arr(d.str_out_dst, d.m_str, np.uint32),
            "layer_off": arr(d.layer_off, d.num_layers + 1, np.uint64), "cc_off": arr(d.cc_off, d.num_cc + 1, np.uint64),
            "rank": arr(d.rank, d.n_s, np.uint32),
            "up_off": arr(d.up_off, d.n_s + 1, np.uint64), "up_v": arr(d.up_v, d.m_up, np.uint32),
            "up_w": arr(d.up_w, d.m_up, np.int32), "dn_off": arr(d.dn_off, d.n_s + 1, np.uint64),
            "dn_v": arr(d.dn_v, d.m_dn, np.uint32), "dn_w": arr(d.dn_w, d.m_dn, np.int32),
            "up_mid": arr(d.up_mid, d.m_up, np.int32), "dn_mid": arr(d.dn_mid, d.m_dn, np.int32),
        }

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            try:
                N.host_lib().chsga_host_free(self._h)
            except Exception:
                pass
            self._h = C.c_void_p()


class ReadSet:
    """Reads as one concatenated byte string plus offsets."""

    def __init__(self, data: bytes, offsets: np.ndarray):
        self.data = data
        self.offsets = np.asarray(offsets, dtype=np.int64)

    @classmethod
    def _from_handle(cls, h: C.c_void_p) -> "ReadSet":
        lib = N.host_lib()
        n = lib.chsga_reads_count(h)
        off = np.ctypeslib.as_array(C.cast(lib.chsga_reads_offsets(h), N.i64p), shape=(n + 1,)).copy()
        data = C.string_at(lib.chsga_reads_data(h), int(off[-1]))
        lib.chsga_reads_free(h)
        return cls(data, off)

    @classmethod
    def generate(cls, ds: Dataset, seed: int, n_reads: int, length: float, len_mode: int = 0, len_sigma: float = 0.0,
                 len_min: int = 1, len_max: int = 1 << 40, sub: float = 0.0, ins: float = 0.0, dele: float = 0.0,
                 n_rate: float = 0.0, mix_frac: float = 0.0, len2=(1, 1), err_scale=(1.0, 1.0)) -> "ReadSet":
        p = N.ReadParams(seed, n_reads, len_mode, float(length), float(len_sigma), len_min, len_max, sub, ins, dele,
                         n_rate, mix_frac, len2[0], len2[1], err_scale[0], err_scale[1])
        h = C.c_void_p()
        N._check_host(N.host_lib().chsga_host_gen_reads(ds._h, C.byref(p), C.byref(h)))
        return cls._from_handle(h)

    @classmethod
    def load(cls, path: str, cap: int = 0) -> "ReadSet":
        h = C.c_void_p()
        N._check_host(N.host_lib().chsga_host_load_reads(str(path).encode(), cap, C.byref(h)))
        return cls._from_handle(h)

    @classmethod
    def from_list(cls, reads: list) -> "ReadSet":
        bs = [r.encode() if isinstance(r, str) else bytes(r) for r in reads]
        off = np.zeros(len(bs) + 1, np.int64)
        off[1:] = np.cumsum([len(b) for b in bs])
        return cls(b"".join(bs), off)

    def __len__(self) -> int:
        return len(self.offsets) - 1

    def __getitem__(self, i: int) -> bytes:
        return self.data[self.offsets[i]:self.offsets[i + 1]]

    def lengths(self) -> np.ndarray:
        return np.diff(self.offsets)

    def subset(self, idx) -> "ReadSet":
        return ReadSet.from_list([self[int(i)] for i in idx])

    def ptrs(self):
        """(char** array, int64 lengths) views for the C ABI; keep the result alive."""
        buf = C.create_string_buffer(self.data, len(self.data) + 1)
        base = C.addressof(buf)
        n = len(self)
        arr = (C.c_char_p * max(n, 1))()
        for i in range(n):
            arr[i] = C.cast(base + int(self.offsets[i]), C.c_char_p)
        lens = np.ascontiguousarray(self.lengths(), dtype=np.int64)
        return buf, arr, lens


@dataclass
class TracebackResult:
    score: int
    end_vertex: int
    ops: np.ndarray  # structured: row int64, vertex uint32, op uint8

    def cigar(self) -> str:
        """ops_to_cigar (src/traceback.hpp:1081-1110): D / M / I run lengths."""
        if len(self.ops) == 0:
            return ""
        ch = {N.OP_DELETE: "D", N.OP_MATCHSUB: "M"}
        out, prev, run = [], None, 0
        for o in self.ops["op"]:
            c = ch.get(int(o), "I")
            if c == prev:
                run += 1
            else:
                if prev is not None:
                    out.append(f"{run}{prev}")
                prev, run = c, 1
        out.append(f"{run}{prev}")
        return "".join(out)


OP_DTYPE = np.dtype([("row", np.int64), ("vertex", np.uint32), ("op", np.uint8)], align=True)


def verify_path(desc: N.GraphDesc, read: bytes, ops: np.ndarray, score: int) -> tuple:
    """The drivers' path verification (src/ConSGA_traceback.cpp:184-225,
    chsga_host_verify_path): (0, -1) when the path follows char edges and
    rescores to `score` with its last consuming row at len - 1, else
    (code, index of the failing op)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    bad = C.c_int64(-1)
    rc = N.host_lib().chsga_host_verify_path(C.byref(desc), read, len(read),
                                             ops.ctypes.data_as(C.POINTER(N.AlignOp)), len(ops), int(score),
                                             C.byref(bad))
    if rc < 0:
        N._check_host(rc)
    return rc, bad.value


class Engine:
    """A chsga_ctx on one GPU."""

    def __init__(self, ds: Dataset, alphabet: int = N.ALPHABET_MISMATCH_N, device: int = 0, desc=None,
                 packed: bool = True, exact: bool = False, cut: bool = False):
        """desc: a GraphDesc of ds to use instead of ds.desc(alphabet), e.g.
        with other edit costs (the context copies what it needs).  packed:
        let the batched path store rows as 2-bit deltas where that is exact
        (compact graphs, unit costs); False keeps one score per char.
        exact: the exact insertion closure of every row (ParSGA's
        semantics) instead of ConSGA's stale last values.  cut: exact, with
        the plain path closing rows on match-cut string components
        (ParSGA's cut-graph insertion) instead of the CH sweep."""
        self.ds = ds
        self.alphabet = alphabet if desc is None else int(desc.alphabet)
        self._lib = N.engine_lib()
        h = C.c_void_p()
        d = ds.desc(alphabet) if desc is None else desc
        flags = ((0 if packed else N.CREATE_CHAR_ROWS) | (N.CREATE_EXACT_CLOSURE if exact else 0)
                 | (N.CREATE_CUT_CLOSURE if cut else 0))
        N._check_engine(self._lib.chsga_ctx_create_ex(C.byref(d), device, flags, C.byref(h)))
        self._h = h
        self.last_stats: dict = {}

    def set_option(self, opt: int, value: int) -> None:
        N._check_engine(self._lib.chsga_ctx_set_option(self._h, opt, value))

    @property
    def packed(self) -> bool:
        """True when the batched rows are packed (2-bit deltas per char)."""
        return self._lib.chsga_ctx_set_option(self._h, N.OPT_PACKED_ROWS, 1) == 0

    # --- batched schedule and its autotuner ----------------------------------
    def schedule(self) -> dict:
        sc = N.BatchSchedule()
        N._check_engine(self._lib.chsga_get_batch_schedule(self._h, C.byref(sc)))
        return sc.as_dict()

    def set_schedule(self, tile_chars: int, big_level: int, big_vpw: int) -> None:
        sc = N.BatchSchedule(tile_chars, big_level, big_vpw)
        N._check_engine(self._lib.chsga_set_batch_schedule(self._h, C.byref(sc)))

    def autotune(self, reads: "ReadSet", ticks: int = 16) -> list:
        """Times candidate batched schedules on these reads and keeps the
        fastest (chsga_autotune_batch); returns every candidate tried as
        dicts with its microseconds per tick."""
        buf, arr, lens = reads.ptrs()
        res = (N.TuneResult * 64)()
        n = C.c_int32()
        N._check_engine(self._lib.chsga_autotune_batch(self._h, arr, lens.ctypes.data_as(N.i64p), len(reads), ticks,
                                                       res, 64, C.byref(n)))
        return [dict(res[i].sched.as_dict(), us_per_tick=res[i].us_per_tick) for i in range(min(n.value, 64))]

    def align_combo(self, reads: ReadSet, per_read_times: bool = False):
        n = len(reads)
        buf, arr, lens = reads.ptrs()
        scores = np.zeros(n, np.int32)
        st = N.Stats()
        tt = np.zeros(n, np.float64) if per_read_times else None
        ts = np.zeros(n, np.float64) if per_read_times else None
        ti = np.zeros(n, np.float64) if per_read_times else None
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None
        N._check_engine(self._lib.chsga_score_reads(self._h, arr, lens.ctypes.data_as(N.i64p), n,
                                                    scores.ctypes.data_as(N.i32p), dp(tt), dp(ts), dp(ti),
                                                    C.byref(st)))
        self.last_stats = st.as_dict()
        if per_read_times:
            return scores, (tt, ts, ti)
        return scores

    def last_row(self, read: bytes) -> np.ndarray:
        V = self.ds.desc(self.alphabet).V
        row = np.zeros(V, np.int32)
        N._check_engine(self._lib.chsga_last_row(self._h, read, len(read), row.ctypes.data_as(N.i32p)))
        return row

    def score_batch(self, reads: ReadSet, slots: int = 64, order=None, use_graphs: bool = False) -> np.ndarray:
        n = len(reads)
        buf, arr, lens = reads.ptrs()
        scores = np.zeros(n, np.int32)
        ordv = None
        if order is not None:
            ordv = np.ascontiguousarray(order, dtype=np.int64)
        opts = N.BatchOpts(slots, 1 if use_graphs else 0, ordv.ctypes.data_as(N.i64p) if ordv is not None else None)
        st = N.Stats()
        N._check_engine(self._lib.chsga_score_batch(self._h, arr, lens.ctypes.data_as(N.i64p), n, C.byref(opts),
                                                    scores.ctypes.data_as(N.i32p), C.byref(st)))
        self.last_stats = st.as_dict()
        return scores

    def align_read_with_traceback(self, read: bytes, k: int = 500) -> TracebackResult:
        p = C.POINTER(N.Path_)()
        st = N.Stats()
        N._check_engine(self._lib.chsga_traceback_read(self._h, read, len(read), k, C.byref(p), C.byref(st)))
        self.last_stats = st.as_dict()
        try:
            n = p.contents.n_ops
            raw = C.string_at(p.contents.ops, n * C.sizeof(N.AlignOp))
            ops = np.frombuffer(raw, dtype=OP_DTYPE).copy()
            return TracebackResult(int(p.contents.score), int(p.contents.end_vertex), ops)
        finally:
            self._lib.chsga_path_free(p)

    def traceback_batch(self, reads: "ReadSet", k: int = 500, slots: int = 64, order=None, use_graphs: bool = False,
                        budget_bytes: int = 0) -> list:
        """Checkpointed traceback of every read with the batched forward pass
        (chsga_traceback_batch); one TracebackResult per read."""
        n = len(reads)
        buf, arr, lens = reads.ptrs()
        ordv = None if order is None else np.ascontiguousarray(order, dtype=np.int64)
        opts = N.BatchOpts(slots, 1 if use_graphs else 0, ordv.ctypes.data_as(N.i64p) if ordv is not None else None)
        paths = (C.POINTER(N.Path_) * n)()
        st = N.Stats()
        N._check_engine(self._lib.chsga_traceback_batch(self._h, arr, lens.ctypes.data_as(N.i64p), n, k,
                                                        C.byref(opts), budget_bytes, paths, C.byref(st)))
        self.last_stats = st.as_dict()
        out = []
        try:
            for i in range(n):
                p = paths[i]
                m = p.contents.n_ops
                raw = C.string_at(p.contents.ops, m * C.sizeof(N.AlignOp))
                out.append(TracebackResult(int(p.contents.score), int(p.contents.end_vertex),
                                           np.frombuffer(raw, dtype=OP_DTYPE).copy()))
        finally:
            for i in range(n):
                if paths[i]:
                    self._lib.chsga_path_free(paths[i])
        return out

    def mem_info(self) -> tuple:
        """(free, total) device bytes of the context's GPU."""
        fr, to = C.c_int64(), C.c_int64()
        N._check_engine(self._lib.chsga_mem_info(self._h, C.byref(fr), C.byref(to)))
        return int(fr.value), int(to.value)

    def traceback_stream(self, reads: "ReadSet", model=None, budget_bytes: int | None = None,
                         candidates=(2000, 1000, 500, 250, 125, 64, 32, 16)) -> list:
        """Checkpointed traceback of every read, each with its own interval k:
        the largest candidate whose peak device bytes fit the budget (default:
        half the free device memory).  The learned cost model proposes k from
        its predicted bytes; the exact byte formula (learned.traceback_peak_bytes)
        checks the proposal and lowers k while it does not fit, and alone picks
        k when no model is given.  Returns (TracebackResult, k) per read."""
        from . import learned as L
        if budget_bytes is None:
            budget_bytes = self.mem_info()[0] // 2
        V = self.ds.info()["V"]
        out = []
        for i in range(len(reads)):
            r = reads[i]
            k = model.choose_k(r, self.ds, budget_bytes, candidates) if model is not None else candidates[0]
            ks = [c for c in candidates if c <= k] or [candidates[-1]]
            for c in ks:  # the formula is the judge: first candidate at or below k that fits
                k = c
                if L.traceback_peak_bytes(len(r), V, c) <= budget_bytes:
                    break
            out.append((self.align_read_with_traceback(r, k=k), k))
        return out

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.chsga_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()


class PasgalAligner:
    """PaSGAL_match, the reference's baseline local aligner
    (src/PaSGALoverrides/align.hpp), on the GPU: two exact-closure contexts,
    one on the char DAG and one on its reversal (chsga_pasgal_align)."""

    def __init__(self, ds: Dataset, device: int = 0, score_only: bool = False, seed: int = 1):
        self.fwd = Engine(ds, device=device, exact=True)
        self.rev_ds = None if score_only else ds.reverse(seed)
        self.rev = None if score_only else Engine(self.rev_ds, device=device, exact=True)
        self.last_stats: dict = {}

    def align(self, reads: ReadSet, match: int = 1, mismatch: int = 1, ins: int = 1, dele: int = 1,
              score_only: bool = False) -> list:
        """One dict per read: score, strand, qry_start/end, ref_start/end (char
        ids), both strands' phase-1 results, rev_best, flags and the compact
        CIGAR (None when score_only)."""
        n = len(reads)
        buf, arr, lens = reads.ptrs()
        prm = N.PasgalParams(match, mismatch, ins, dele, 1 if score_only else 0)
        res = (N.PasgalResult * max(n, 1))()
        st = N.Stats()
        lib = self.fwd._lib
        rev_h = self.rev._h if self.rev is not None else None
        N._check_engine(lib.chsga_pasgal_align(self.fwd._h, rev_h, C.byref(prm), arr, lens.ctypes.data_as(N.i64p), n,
                                               res, C.byref(st)))
        self.last_stats = st.as_dict()
        try:
            return [res[i].as_dict() for i in range(n)]
        finally:
            lib.chsga_pasgal_free(res, n)

    def close(self) -> None:
        self.fwd.close()
        if self.rev is not None:
            self.rev.close()
