#!/usr/bin/env python3
"""GCUPS benchmark of the ConSGA graph-DP hot path on one MI355X.

A step is one pass of the hot path over one batch of synthetic reads of the
c3 workload (20 Mbp seeded chain-of-bubbles DAG, ~22M char vertices; reads
log-normal median 3 kbp clipped to 1-30 kbp, 2% error, 1% N, mismatchN):
the optimal score and an alignment path of every read, with checkpoint rows
every k = 500 kept in HBM.  Step k takes reads k*B .. (k+1)*B - 1 of the
4096-read c3 stream (cycling), B = --batch-reads.
  value      GCUPS of the batched path with traceback (slot-interleaved rows,
             device work queue, reads packed in order of the learned cost
             model; forward pass keeping every 500th row, then segment
             recompute and walk of 64 reads at a time), whole job, inputs
             resident in HBM; score_only_gcups is its forward pass alone
             (--score-only times score-only batches instead);
  plain      GCUPS of the plain path (one read per launch sequence,
             score-only) on the same graph;
  roofline   the batched row kernel (k_brow_packed on c3; k_brow_tile where
             rows are char rows), one DP row of 64 reads per launch:
             algorithmic bytes per launch over its mean duration from HIP
             events on the engine's stream; traffic = HBM bytes per launch
             from the rocprofv3 FETCH_SIZE/WRITE_SIZE summary in profiles/
             (scaled by a plain copy of known size), when that summary was
             taken on the same per-launch bytes;
  check      after the timed region: every path of every timed step verified
             (chsga_host_verify_path: char edges, rescored cost = score, last
             consuming row = |r| - 1) and the batched scores of a sample of
             reads against the plain path's;
  cpu_baseline  the oracle (C++/OpenMP restatement of ConSGA) on a bounded
             sample of the same workload on this host's cores.

Wall clock: the run keeps to --time-budget seconds (default 540) from its
start.  A timed step is started only while the mean step so far, plus the
time reserved for the legs after the timed region, still fits; `steps`
reports the steps actually timed (`steps_requested` the ones asked for).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GCUPS (read chars x graph chars / s) on 1 GPU: batched+learned vs plain, exact"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def row_kernel_bytes(info: dict, score_bytes: int = 4) -> int:
    """Algorithmic (compulsory) bytes of one compact k_row_tile8 launch (one
    DP row of one read): read H and write D (2 V scores), labels V, the
    first-char bit mask V/8, the 8-byte first_nb entry of every string's first
    char (its first two in-neighbours; the in-CSR is read only for the rare
    first char with more) and the F and Lst writes (2 n_s scores).  The
    gathered in-neighbour scores are lines of the same input row."""
    V, n_s = info["V"], info["n_s"]
    return 2 * score_bytes * V + V + V // 8 + 8 * n_s + 2 * score_bytes * n_s


def batch_kernel_bytes(info: dict, score_bytes: int, slots: int = 64) -> int:
    """Algorithmic (compulsory) bytes of one k_brow_tile launch (one tick: one
    DP row of each of `slots` reads): every byte it must move once — the
    slots' input row read and output row written (2 x slots x score bytes per
    char), one label class byte and one first-char bit per char, and the
    16-byte in-neighbour entry of every string.  In-neighbour rows gathered
    for first chars are lines of the same input row (mostly L2 hits), so
    they add no compulsory bytes; the measured traffic shows what they cost."""
    V, n_s = info["V"], info["n_s"]
    return 2 * slots * score_bytes * V + V + V // 8 + 16 * n_s


def packed_layout(ds, tile_chars: int) -> tuple:
    """(delta word groups, row-pass chunks) of packed rows under the engine's
    tiling: whole strings per tile, at most tile_chars chars unless one
    string is longer; each tile's deltas take whole 16-char groups; chunks
    of at most 32 chars, cut before a tile's fifth string start within
    them."""
    a = ds.arrays()
    so = [int(x) for x in a["str_off"]]
    n_s = len(so) - 1
    groups = chunks = 0
    s = 0
    while s < n_s:  # one tile: strings s .. e - 1
        e, acc = s, 0
        while e < n_s and (acc == 0 or acc + so[e + 1] - so[e] <= tile_chars):
            acc += so[e + 1] - so[e]
            e += 1
        groups += (acc + 15) // 16
        # chunks: string starts of the tile, relative to its first char
        c0, c1 = so[s], so[e]
        starts = [so[i] - c0 for i in range(s, e)]
        p, i, n = 0, 0, c1 - c0
        while p < n:
            while i < len(starts) and starts[i] < p:
                i += 1
            cut = starts[i + 4] - p if i + 4 < len(starts) and starts[i + 4] - p < 64 else 64
            p += min(32, cut, n - p)
            chunks += 1
        s = e
    return groups, chunks


def packed_kernel_bytes(ds, tile_chars: int, score_bytes: int, slots: int = 64) -> int:
    """Algorithmic (compulsory) bytes of one k_brow_packed launch (packed
    rows: 2-bit codes per char and slot, F/PL/PF planes per string).  Read
    and written once each: the delta words of every tile (whole 16-char
    groups per tile, 4 bytes per slot), plane F of every string and planes
    PF and PL of every string of more than one char (score bytes per slot);
    read once: the 128-byte table entry of every row-pass chunk (char count,
    string starts, label classes, gathered planes).  Gathered in-neighbour
    planes are the same planes of the input row."""
    a = ds.arrays()
    lens = np.diff(a["str_off"].astype(np.int64))
    groups, chunks = packed_layout(ds, tile_chars)
    n_s, n_multi = len(lens), int((lens > 1).sum())
    return 2 * groups * slots * 4 + 2 * (n_s + 2 * n_multi) * slots * score_bytes + 128 * chunks


def profile_summary(kernel: str, bytes_per_launch: int) -> dict:
    """Figures of the committed counter summary of `kernel` on c3 (HBM
    traffic, VALU instructions per launch), or {} when it was taken on
    different algorithmic bytes per launch."""
    p = ROOT / "profiles" / f"{kernel}_c3_hbm_traffic.md"
    if not p.exists():
        return {}
    vals = {}
    for line in p.read_text().splitlines():
        k, _, v = line.partition(":")
        if v.strip().isdigit():
            vals[k.strip()] = int(v.strip())
    if vals.get("algorithmic_bytes_per_launch") != bytes_per_launch:
        return {}
    return vals


# VALU issue rate of MI355X: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64
# instruction per SIMD every 4 cycles (16 lanes); SALU: one instruction per
# CU per cycle
VALU_PEAK_INSTS = 256 * 4 * 2.4e9 / 4
SALU_PEAK_INSTS = 256 * 2.4e9


def row_bytes_survey(info: dict) -> int:
    """Whole-row algorithmic bytes: 4V + 4V + V + 4(V+1) + 4E + 12 m_CH + 8 n_s."""
    V, E, n_s = info["V"], info["E"], info["n_s"]
    return 4 * V + 4 * V + V + 4 * (V + 1) + 4 * E + 12 * (info["m_up"] + info["m_dn"]) + 8 * n_s


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ds, wl, reads, budget_s: float, gpu_engine) -> dict:
    from oracle import oracle as O
    threads = O.threads(0)
    d = ds.desc(wl.alphabet)
    V = d.V
    # bounded sample: a prefix of read 0, sized from a short probe
    probe_rows = 16
    r0 = reads[0]
    t = time.perf_counter()
    O.consga_scores(d, type(reads).from_list([r0[:probe_rows]]))
    per_row = (time.perf_counter() - t) / probe_rows
    rows = int(max(32, min(len(r0), budget_s / max(per_row, 1e-6))))
    sample = type(reads).from_list([r0[:rows]])
    t = time.perf_counter()
    cpu_score = O.consga_scores(d, sample)
    dt = time.perf_counter() - t
    gpu_score = gpu_engine.align_combo(sample)
    return {
        "value": round(rows * V / dt / 1e9, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
        "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
        "sample": f"score-only DP of read 0 of the c3 read set truncated to its first {rows} chars "
                  f"({rows} DP rows x {V} chars; paths are not timed on the CPU), "
                  f"{dt:.1f} s on {threads} OpenMP threads",
        "score_cpu": int(cpu_score[0]), "score_gpu": int(gpu_score[0]), "exact": bool(cpu_score[0] == gpu_score[0]),
    }


def pasgal_leg(ds, stream, args, no_cpu: bool) -> dict:
    """PaSGAL_match (the reference's baseline local aligner) on the same
    graph: GCUPS of phase 1 alone (both strands) and of the whole alignment
    (phases 1, 1-R and 2 with CIGARs), each over the DP cells it computes,
    on --pasgal-reads reads cut to --pasgal-len chars; and its CPU
    restatement on a short prefix of read 0, whose results must equal the
    GPU's."""
    from chsga_amd.api import PasgalAligner
    reads = type(stream).from_list([stream[i][:args.pasgal_len] for i in range(args.pasgal_reads)])
    t = time.perf_counter()
    al = PasgalAligner(ds)
    setup = time.perf_counter() - t
    out = {"reads": len(reads), "read_chars": args.pasgal_len, "context_setup_s": round(setup, 1)}
    for mode, so in (("phase1", True), ("full", False)):
        t = time.perf_counter()
        res = al.align(reads, score_only=so)
        dt = time.perf_counter() - t
        out[f"{mode}_gcups"] = round(al.last_stats["cells"] / dt / 1e9, 3)
        out[f"{mode}_s"] = round(dt, 2)
    out["scores"] = [r["score"] for r in res]
    if not no_cpu:
        from oracle import oracle as O
        sample = type(stream).from_list([stream[0][:24]])
        t = time.perf_counter()
        exp = O.pasgal_align(ds.desc(), sample)
        dt = time.perf_counter() - t
        V = ds.info()["V"]
        e = exp[0]
        window = max(0, e["qry_end"] - e["qry_start"] + 1) * max(0, e["ref_end"] - e["ref_start"] + 1)
        cells = 3 * len(sample[0]) * V + window  # phases 1 (two strands) and 1-R, phase 2 on its window
        got = al.align(sample)
        out["cpu"] = {"gcups": round(cells / dt / 1e9, 4), "cores": 1, "kind": "port",
                      "sample": f"read 0 cut to {len(sample[0])} chars, whole alignment, one thread",
                      "exact": got == exp}
    al.close()
    return out


def main() -> int:
    t_start = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--batch-reads", type=int, default=1024, help="reads per batched step")
    ap.add_argument("--k", type=int, default=500, help="checkpoint interval of the traceback")
    ap.add_argument("--score-only", action="store_true", help="batched steps compute scores only (no paths)")
    ap.add_argument("--plain-reads", type=int, default=2, help="reads per plain step")
    ap.add_argument("--plain-budget", type=float, default=40.0, help="seconds for the plain leg's timed steps")
    ap.add_argument("--slots", type=int, default=128,
                    help="batched slots: 64 per slot group, each group on its own stream with its own backward worker")
    ap.add_argument("--mode", choices=["both", "plain", "batch"], default="both")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--time-budget", type=float, default=540.0,
                    help="wall-clock seconds for the whole run (timed steps stop early to keep it)")
    ap.add_argument("--check-reads", type=int, default=4,
                    help="reads of the last timed step whose batched scores are checked against the plain path")
    ap.add_argument("--forward-tail", type=int, default=None,
                    help="CHSGA_OPT_FORWARD_TAIL (busy slots on average below which a batch's tail runs on the plain path)")
    ap.add_argument("--backward-tail", type=int, default=None,
                    help="CHSGA_OPT_BACKWARD_TAIL (ratio below which the backward's last reads go read by read)")
    ap.add_argument("--first-step", type=int, default=0, help="index of the first timed step's batch in the stream")
    ap.add_argument("--calibrate-copy", type=int, default=0, help="only run a torch copy of N MiB (PMC scaling)")
    ap.add_argument("--traceback", type=int, default=0,
                    help="also time the per-read checkpointed traceback of this many reads")
    ap.add_argument("--pasgal-reads", type=int, default=2,
                    help="reads (cut to --pasgal-len chars) of the PaSGAL_match comparison leg (0: none)")
    ap.add_argument("--pasgal-len", type=int, default=1000)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the path shards by reads and no multi-GPU figure is taken "
              "(use --gpus 1)", file=sys.stderr)
        return 2

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy} MiB x5")
        return 0

    from chsga_amd import WORKLOADS, Engine
    from chsga_amd import learned as L
    from chsga_amd import native as N
    from chsga_amd.api import verify_path
    from chsga_amd.scheduler import plan_order

    def elapsed() -> float:
        return time.perf_counter() - t_start

    wl = WORKLOADS[args.workload]
    t = time.perf_counter()
    ds = wl.dataset()
    info = ds.info()
    log(f"[bench] {wl.name} graph: V={info['V']} E={info['E']} n_s={info['n_s']} levels={info['num_layers']} "
        f"({time.perf_counter() - t:.1f}s)")
    stream = wl.readset(ds)  # the workload's read stream; steps take consecutive batches, cycling
    n_stream = len(stream)
    eng = Engine(ds, wl.alphabet)
    eng.set_option(N.OPT_PROFILE, 0)
    if args.forward_tail is not None:
        eng.set_option(N.OPT_FORWARD_TAIL, args.forward_tail)
    if args.backward_tail is not None:
        eng.set_option(N.OPT_BACKWARD_TAIL, args.backward_tail)
    # queue order: longest predicted time first, predictions from the learned
    # cost model (trained on the GPU, chsga_amd/data) or the plain formula
    model = L.load_default()
    scheduler = "learned cost model (LPT)" if model is not None else "plain formula rows x V (LPT)"
    log(f"[bench] scheduler: {scheduler}")
    V = info["V"]
    desc = ds.desc(wl.alphabet)
    result = {}

    def step_reads(k: int, per: int):
        return stream.subset([(k * per + j) % n_stream for j in range(per)])

    # time reserved for the legs after the batched timed region: plain leg,
    # roofline profile, path checks, CPU baseline
    reserve = (args.plain_budget + 15.0 if args.mode in ("both", "plain") else 0.0) + 25.0 + \
        (0.0 if args.no_cpu else args.cpu_budget + 15.0) + (45.0 if args.pasgal_reads > 0 else 0.0)

    # ---- batched path -------------------------------------------------------
    # A step is one batch of batch_reads reads through the 64 slots: scores
    # and paths (checkpointed traceback), or scores only with --score-only;
    # the warm-up is a short batch that allocates the slot buffers.
    batch_gcups = score_gcups = None
    kept = []  # (reads, scores or paths) of every timed step, checked after the timed region
    steps_done = 0
    if args.mode in ("both", "batch"):
        for w in range(args.warmup):
            rs = step_reads(w, args.batch_reads)
            rs = type(rs).from_list([rs[i][:64] for i in range(min(len(rs), 2 * args.slots))])
            eng.score_batch(rs, slots=args.slots, order=plan_order(rs, ds, model))
            if not args.score_only:
                eng.traceback_batch(rs, k=args.k, order=plan_order(rs, ds, model))
            log(f"[bench] batched warm-up {w} done ({elapsed():.0f}s)")
        # batched schedule from the autotuner (setup, outside the timed region)
        tried = eng.autotune(step_reads(0, args.batch_reads), ticks=24)
        best = min(tried, key=lambda t: t["us_per_tick"])
        log(f"[bench] autotuned schedule {eng.schedule()} ({best['us_per_tick']:.0f} us/tick; "
            f"{len(tried)} candidates, {tried[0]['us_per_tick']:.0f} us/tick for the default)")
        result["schedule"] = eng.schedule()
        result["autotune"] = [{k: (round(v, 1) if k == "us_per_tick" else v) for k, v in t.items()} for t in tried]
        cells = ticks = fallbacks = ops = peak = 0
        fwd_s = 0.0
        t0 = time.perf_counter()
        for k in range(args.steps):
            if k > 0:
                mean = (time.perf_counter() - t0) / k
                if elapsed() + mean + reserve > args.time_budget:
                    log(f"[bench] wall-clock budget: stopping after {k} of {args.steps} timed steps "
                        f"({elapsed():.0f}s elapsed, {mean:.1f}s per step, {reserve:.0f}s reserved)")
                    break
            rs = step_reads(args.warmup + args.first_step + k, args.batch_reads)
            ts = time.perf_counter()
            order = plan_order(rs, ds, model)
            if args.score_only:
                sc = eng.score_batch(rs, slots=args.slots, order=order, use_graphs=True)
                fwd_s += time.perf_counter() - ts
                kept.append((rs, sc))
            else:
                paths = eng.traceback_batch(rs, k=args.k, slots=args.slots, order=order, use_graphs=True)
                fwd_s += eng.last_stats["forward_ms"] / 1e3
                ops += sum(len(p.ops) for p in paths)
                kept.append((rs, paths))
            st = eng.last_stats
            cells += int(rs.lengths().sum()) * V
            ticks += st["rows"]
            fallbacks += st["walk_fallbacks"]
            peak = max(peak, st["peak_device_bytes"])
            steps_done += 1
            log(f"[bench] batched step {k}: {len(rs)} reads, {st['rows']} ticks, {time.perf_counter() - ts:.1f}s "
                f"({elapsed():.0f}s elapsed)")
        dt = time.perf_counter() - t0
        batch_gcups = cells / dt / 1e9
        score_gcups = cells / fwd_s / 1e9
        result["batch"] = {"gcups": round(batch_gcups, 3), "ms_per_step": round(dt * 1e3 / steps_done, 2),
                           "steps": steps_done, "reads_per_step": args.batch_reads, "slots": args.slots,
                           "work": "scores only" if args.score_only else
                           f"scores and paths (checkpoint rows every {args.k} in HBM)",
                           "score_only_gcups": round(score_gcups, 3), "ticks": ticks}
        if not args.score_only:
            result["batch"].update({"path_ops": ops, "walk_fallbacks": fallbacks,
                                    "peak_device_gb": round(peak / 1e9, 1)})
        log(f"[bench] batched: {batch_gcups:.2f} GCUPS ({dt:.2f}s, {steps_done} steps); "
            f"forward pass alone {score_gcups:.2f} GCUPS")

    # ---- plain path: one read per launch sequence -------------------------
    plain_gcups = None
    plain_steps = 0
    if args.mode in ("both", "plain"):
        for w in range(min(args.warmup, 1)):
            eng.align_combo(step_reads(w, args.plain_reads))
            log(f"[bench] plain warm-up {w} done")
        cells = 0
        t0 = time.perf_counter()
        for k in range(args.steps):
            if k > 0 and (time.perf_counter() - t0) * (k + 1) / k > args.plain_budget:
                break
            rs = step_reads(args.warmup + k, args.plain_reads)
            eng.align_combo(rs)
            cells += int(rs.lengths().sum()) * V
            plain_steps += 1
        dt = time.perf_counter() - t0
        plain_gcups = cells / dt / 1e9
        result["plain"] = {"gcups": round(plain_gcups, 3), "ms_per_step": round(dt * 1e3 / plain_steps, 2),
                           "steps": plain_steps, "reads_per_step": args.plain_reads}
        log(f"[bench] plain: {plain_gcups:.2f} GCUPS ({dt:.2f}s, {plain_steps} steps; {elapsed():.0f}s elapsed)")

    # ---- roofline of the dominant kernel (profiled, separate short run) ----
    # HIP events bracket every row-kernel launch on the engine's own stream;
    # one slot group (64 slots), so that no other group's kernels overlap the
    # launches timed
    score_bytes = 2 if max(stream.lengths()) < 32767 else 4  # int16 storage when provably exact
    eng.set_option(N.OPT_PROFILE, 1)
    prof_slots = 64
    prof_b = stream.subset(range(min(n_stream, 3 * prof_slots)))
    prof_b = type(prof_b).from_list([prof_b[i][:400] for i in range(len(prof_b))])
    eng.score_batch(prof_b, slots=prof_slots)
    stb = eng.last_stats
    eng.align_combo(stream.subset([0]))
    stp = eng.last_stats
    eng.set_option(N.OPT_PROFILE, 0)
    b_ms = stb["row_kernel_ms"] / max(1, stb["row_kernel_launches"])
    p_ms = stp["row_kernel_ms"] / max(1, stp["row_kernel_launches"])
    packed = eng.packed
    kernel = "k_brow_packed" if packed else "k_brow_tile"
    kb = (packed_kernel_bytes(ds, eng.schedule()["tile_chars"], score_bytes, prof_slots) if packed
          else batch_kernel_bytes(info, score_bytes, prof_slots))
    pb = row_kernel_bytes(info, score_bytes)
    achieved = kb / (b_ms * 1e-3) / 1e9
    prof = profile_summary(kernel, kb)
    valu = salu = None
    if prof.get("valu_insts_per_launch"):
        at_peak = prof["valu_insts_per_launch"] / VALU_PEAK_INSTS
        valu = {"insts_per_launch": prof["valu_insts_per_launch"], "peak_insts_per_s": VALU_PEAK_INSTS,
                "us_at_peak": round(at_peak * 1e6, 1), "frac": round(at_peak / (b_ms * 1e-3), 4)}
    if prof.get("salu_insts_per_launch"):
        at_peak = prof["salu_insts_per_launch"] / SALU_PEAK_INSTS
        salu = {"insts_per_launch": prof["salu_insts_per_launch"], "peak_insts_per_s": SALU_PEAK_INSTS,
                "us_at_peak": round(at_peak * 1e6, 1), "frac": round(at_peak / (b_ms * 1e-3), 4)}
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": prof.get("traffic_bytes_per_launch"),
                "valu": valu, "salu": salu, "kernel": kernel, "kernel_avg_us": round(b_ms * 1e3, 2),
                "bytes_per_launch": kb, "launches_timed": stb["row_kernel_launches"],
                "score_storage_bytes": score_bytes, "slots_profiled": prof_slots,
                "plain_kernel": {"kernel": "k_row_tile8", "avg_us": round(p_ms * 1e3, 2), "bytes_per_launch": pb,
                                 "achieved_GBps": round(pb / (p_ms * 1e-3) / 1e9, 1),
                                 "sweep_post_avg_us_per_row": round(stp["sweep_ms"] * 1e3 /
                                                                    max(1, stp["row_kernel_launches"]), 2)},
                "batched_sweep_post_avg_us_per_tick": round(stb["sweep_ms"] * 1e3 /
                                                            max(1, stb["row_kernel_launches"]), 2)}
    log(f"[bench] {kernel}: {b_ms * 1e3:.1f} us/launch, {achieved:.0f} GB/s; "
        f"k_row_tile8: {p_ms * 1e3:.1f} us/launch")

    # ---- checks of the timed work (outside the timed region) ---------------
    check = None
    if kept:
        t1 = time.perf_counter()
        n_paths = bad_paths = 0
        first_bad = None
        for rs, res in kept:
            if args.score_only:
                break
            for i, p in enumerate(res):
                rc, at = verify_path(desc, rs[i], p.ops, p.score)
                n_paths += 1
                if rc != 0:
                    bad_paths += 1
                    first_bad = first_bad or {"read": i, "code": rc, "op": at}
        rs, res = kept[-1]
        sc = np.array([p.score for p in res] if not args.score_only else res, dtype=np.int64)
        # a sample of the last step's reads: the shortest ones and an evenly spread few
        lens = rs.lengths()
        pick = sorted(set(np.argsort(lens, kind="stable")[:max(1, args.check_reads // 2)].tolist() +
                          np.linspace(0, len(rs) - 1, max(1, args.check_reads - args.check_reads // 2),
                                      dtype=int).tolist()))
        plain = eng.align_combo(rs.subset(pick)).astype(np.int64)
        check = {"paths_verified": n_paths, "paths_failed": bad_paths, "first_failure": first_bad,
                 "scores_checked_against_plain": len(pick), "read_lengths_checked": lens[pick].tolist(),
                 "scores_equal_plain": bool(np.array_equal(plain, sc[pick])),
                 "scores_in_bounds": bool(((sc >= 0) & (sc <= lens)).all()),
                 "seconds": round(time.perf_counter() - t1, 1)}
        log(f"[bench] checks: {check}")

    # ---- optional: the per-read checkpointed traceback, for comparison ----
    if args.traceback > 0:
        rs = step_reads(0, args.traceback)
        t1 = time.perf_counter()
        for i in range(len(rs)):
            eng.align_read_with_traceback(rs[i], k=args.k)
        dt1 = time.perf_counter() - t1
        cells1 = int(rs.lengths().sum()) * V
        result["traceback_per_read"] = {"reads": len(rs), "k": args.k, "gcups": round(cells1 / dt1 / 1e9, 3)}
        log(f"[bench] per-read traceback: {cells1 / dt1 / 1e9:.2f} GCUPS ({len(rs)} reads)")

    # ---- PaSGAL_match, the reference's baseline aligner, for comparison ----
    if args.pasgal_reads > 0:
        result["pasgal"] = pasgal_leg(ds, stream, args, args.no_cpu)
        log(f"[bench] PaSGAL_match: {result['pasgal']}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ds, wl, stream, args.cpu_budget, eng)
        log(f"[bench] cpu baseline: {cpu['value']} GCUPS on {cpu['cores']} threads ({cpu['cpu_model']}), "
            f"exact={cpu['exact']}")

    value = batch_gcups if batch_gcups is not None else plain_gcups
    main_leg = result.get("batch", result.get("plain"))
    n_reads = args.batch_reads if batch_gcups is not None else args.plain_reads
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "GCUPS", "n_gpus": 1, "steps": main_leg["steps"],
        "steps_requested": args.steps, "warmup": args.warmup, "ms_per_step": main_leg["ms_per_step"],
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded graph and reads)",
        "config": {"workload": f"{n_reads} reads of the {wl.name} stream per step: 20 Mbp seeded DAG (SNP 1/25, "
                               f"indels 1-50 at 1/200), V={V} chars, n_s={info['n_s']}; reads log-normal median "
                               f"3 kbp (1-30 kbp), 2% error, 1% N, mismatchN, " +
                               ("scores only" if args.score_only or batch_gcups is None else
                                f"scores and paths, checkpoint rows every {args.k} in HBM"),
                   "batch_reads": args.batch_reads, "plain_reads": args.plain_reads, "slots": args.slots,
                   "scheduler": scheduler, "schedule": result.get("schedule")},
        "autotune": result.get("autotune"),
        "plain_gcups": None if plain_gcups is None else round(plain_gcups, 3),
        "batched_gcups": None if batch_gcups is None else round(batch_gcups, 3),
        "score_only_gcups": None if score_gcups is None else round(score_gcups, 3),
        "speedup_batched_vs_plain": (round(batch_gcups / plain_gcups, 2) if batch_gcups and plain_gcups else None),
        "batch": result.get("batch"), "plain": result.get("plain"), "check": check,
        "roofline": roofline, "cpu_baseline": cpu, "wall_s": round(elapsed(), 1),
    }
    if "traceback_per_read" in result:
        line["traceback_per_read"] = result["traceback_per_read"]
    if "pasgal" in result:
        line["pasgal_match"] = result["pasgal"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
