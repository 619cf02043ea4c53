"""Write a seeded workload as drop-in input files.

    python -m chsga_amd.gen c1 /tmp/c1 [--reads N]

produces /tmp/c1.char.adj (-I), .str.adj (-i), .ch (-c), .labels (-l),
.label_offsets (-o) and .reads.fq (-r) for the ConSGA_* drivers.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


def main() -> int:
    from chsga_amd import WORKLOADS
    from chsga_amd import native as N
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=sorted(WORKLOADS))
    ap.add_argument("prefix")
    ap.add_argument("--reads", type=int, default=None)
    args = ap.parse_args()
    w = WORKLOADS[args.workload]
    ds = w.dataset()
    ds.save(args.prefix)
    rs = w.readset(ds, n_reads=args.reads)
    import ctypes as C
    lib = N.host_lib()
    h = C.c_void_p()
    p = Path(args.prefix + ".reads.fq")
    with open(p, "wb") as f:
        for i in range(len(rs)):
            r = rs[i]
            f.write(b"@read_%d\n%s\n+\n%s\n" % (i, r, b"I" * len(r)))
    print(f"{args.workload}: {ds.info()} -> {args.prefix}.*, {len(rs)} reads")
    return 0


if __name__ == "__main__":
    sys.exit(main())
