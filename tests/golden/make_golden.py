"""Writes the golden fixtures: seeded graphs + reads with the oracle's scores.

The reference ships no data or expected outputs for this path, so these
fixtures are produced by this build's oracle (oracle/oracle.cpp) and pinned
by the oracle's own tests (closed forms, brute force, invariants).  They are
regression vectors for the GPU engine and for the oracle itself.

    python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from chsga_amd import ALPHABET_MATCH_N, ALPHABET_MISMATCH_N, Dataset, ReadSet  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = [
    ("snp_indel_mismatchN", dict(seed=21, backbone_bp=2500, snp_rate=1 / 20, indel_rate=1 / 80, indel_len=(1, 15)),
     dict(seed=1, n_reads=12, length=80, sub=0.03, ins=0.01, dele=0.01, n_rate=0.03), ALPHABET_MISMATCH_N),
    ("snp_indel_matchN", dict(seed=22, backbone_bp=2500, snp_rate=1 / 20, indel_rate=1 / 80, indel_len=(1, 15)),
     dict(seed=2, n_reads=12, length=80, sub=0.03, ins=0.01, dele=0.01, n_rate=0.05), ALPHABET_MATCH_N),
    ("structural", dict(seed=23, backbone_bp=6000, snp_rate=1 / 40, indel_rate=1 / 150, indel_len=(1, 30),
                        sv_rate=1 / 1500, sv_len=(50, 400)),
     dict(seed=3, n_reads=10, length=0, len_mode=2, len_min=5, len_max=40, mix_frac=0.5, len2=(100, 250),
          sub=0.03, ins=0.015, dele=0.015, err_scale=(0.2, 1.0)), ALPHABET_MISMATCH_N),
]


def main():
    for name, gp, rp, alpha in CASES:
        ds = Dataset.generate(**gp)
        rs = ReadSet.generate(ds, **rp)
        d = ds.desc(alpha)
        scores = [int(x) for x in O.consga_scores(d, rs)]
        row = O.consga_last_row(d, rs[0])
        out = {
            "graph": {k: (list(v) if isinstance(v, tuple) else v) for k, v in gp.items()},
            "alphabet": alpha,
            "info": ds.info(),
            "reads": [rs[i].decode() for i in range(len(rs))],
            "scores": scores,
            "exact_scores": [int(x) for x in O.exact_scores(d, rs)],
            "row0_sum": int(row.sum()),
            "row0_argmin": int(row.argmin()),
        }
        (HERE / f"{name}.json").write_text(json.dumps(out, indent=1) + "\n")
        print(name, scores)


if __name__ == "__main__":
    main()
