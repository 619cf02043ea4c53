"""A short fixed workload for profilers (rocprofv3 --kernel-trace / --pmc).

    python -m chsga_amd.profile_run [--workload c3] [--rows 300] [--reads 64] [--slots 64]
                                    [--mode both|plain|batch|traceback] [--k 50]
                                    [--closure consga|exact|cut]

Takes the first `rows` chars of `reads` reads of the workload and runs read 0
on the plain path (rows - 1 DP rows) and all of them on the batched path
(`rows` ticks of `slots` slots, in groups of 64), printing one line per stage.
`--mode traceback` instead runs the batched traceback of those reads with
checkpoint rows every `k` rows: every lane stores one on the same ticks
(rows 0, k, 2k, ...), the most a capture tick can hold; with more reads than
slots the backward's banded rounds each take 64 reads in graph order, so
their ticks stop at increasing chars.  `--closure` picks
the rows' insertion closure: ConSGA's (default), the exact closure by the CH
sweep (CHSGA_CREATE_EXACT_CLOSURE) or by match-cut components on the plain
path (CHSGA_CREATE_CUT_CLOSURE).
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


def main() -> int:
    from chsga_amd import WORKLOADS, Engine, ReadSet
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--rows", type=int, default=300)
    ap.add_argument("--reads", type=int, default=64)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--mode", choices=["both", "plain", "batch", "traceback"], default="both")
    ap.add_argument("--k", type=int, default=50, help="checkpoint interval of --mode traceback")
    ap.add_argument("--closure", choices=["consga", "exact", "cut"], default="consga")
    a = ap.parse_args()
    w = WORKLOADS[a.workload]
    ds = w.dataset()
    rs = w.readset(ds, n_reads=a.reads)
    short = ReadSet.from_list([rs[i][: a.rows] for i in range(len(rs))])
    eng = Engine(ds, w.alphabet, exact=a.closure == "exact", cut=a.closure == "cut")
    print(f"[profile_run] {a.workload}: {ds.info()}", flush=True)
    if a.mode in ("both", "plain"):
        t = time.time()
        s = eng.align_combo(short.subset([0]))
        print(f"[profile_run] plain: 1 read x {len(short[0])} rows in {time.time() - t:.2f}s, score {s[0]}", flush=True)
    if a.mode in ("both", "batch"):
        t = time.time()
        s = eng.score_batch(short, slots=a.slots)
        print(f"[profile_run] batched: {len(short)} reads x {a.rows} rows in {time.time() - t:.2f}s", flush=True)
    if a.mode == "traceback":
        t = time.time()
        paths = eng.traceback_batch(short, k=a.k, slots=a.slots)
        st = eng.last_stats
        print(f"[profile_run] traceback: {len(short)} reads x {a.rows} rows, k = {a.k}, {st['rows']} ticks, "
              f"{sum(len(p.ops) for p in paths)} path ops in {time.time() - t:.2f}s", flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
