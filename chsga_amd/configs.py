"""The five seeded workloads (graph + reads) the measurements and tests use.

Sizes follow the configurations the metric is quoted on:
  c1  100 kbp backbone, SNP 1/50 bp, indels 1-10 bp at 1/500 bp; 64 x 1 kbp reads,
      1% substitutions + 0.5% indels, no N; score-only + traceback k=500.
  c2  5 Mbp, SNP 1/30, indels 1-20 at 1/300; 1000 x 10 kbp reads, 5% error (2:1:1).
  c3  20 Mbp, SNP 1/25, indels 1-50 at 1/200; 4096 reads, log-normal median 3 kbp,
      clipped to 1-30 kbp, 2% error, 1% N, mismatchN; traceback k=500.
  c4  the c2 graph; a stream of 16,384 reads, log-normal 0.5-30 kbp.
  c5  100 Mbp, SNP 1/20, indels 1-50 at 1/150, structural bubbles 50 bp-5 kbp at
      1/20 kbp, matchN; 10,000 reads, half 100-300 bp at 0.5%, half 5-50 kbp at 5%.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from . import native as N
from .api import Dataset, ReadSet


@dataclass
class Workload:
    name: str
    graph: dict
    reads: dict
    alphabet: int = N.ALPHABET_MISMATCH_N
    traceback_k: int = 0
    notes: str = ""
    extra: dict = field(default_factory=dict)

    def dataset(self) -> Dataset:
        return Dataset.generate(**self.graph)

    def readset(self, ds: Dataset, n_reads: int | None = None, seed: int | None = None) -> ReadSet:
        kw = dict(self.reads)
        if n_reads is not None:
            kw["n_reads"] = n_reads
        if seed is not None:
            kw["seed"] = seed
        return ReadSet.generate(ds, **kw)


LOGNORMAL_SIGMA = 0.6  # recorded sigma of the log-normal read lengths

WORKLOADS = {
    "c1": Workload(
        "c1", dict(seed=1, backbone_bp=100_000, snp_rate=1 / 50, indel_rate=1 / 500, indel_len=(1, 10)),
        dict(seed=101, n_reads=64, length=1000, sub=0.01, ins=0.0025, dele=0.0025),
        traceback_k=500),
    "c2": Workload(
        "c2", dict(seed=2, backbone_bp=5_000_000, snp_rate=1 / 30, indel_rate=1 / 300, indel_len=(1, 20)),
        dict(seed=202, n_reads=1000, length=10_000, sub=0.025, ins=0.0125, dele=0.0125)),
    "c3": Workload(
        "c3", dict(seed=3, backbone_bp=20_000_000, snp_rate=1 / 25, indel_rate=1 / 200, indel_len=(1, 50)),
        dict(seed=303, n_reads=4096, length=3000, len_mode=1, len_sigma=LOGNORMAL_SIGMA, len_min=1000,
             len_max=30_000, sub=0.01, ins=0.005, dele=0.005, n_rate=0.01),
        traceback_k=500),
    "c4": Workload(
        "c4", dict(seed=2, backbone_bp=5_000_000, snp_rate=1 / 30, indel_rate=1 / 300, indel_len=(1, 20)),
        dict(seed=404, n_reads=16_384, length=3000, len_mode=1, len_sigma=0.9, len_min=500, len_max=30_000,
             sub=0.025, ins=0.0125, dele=0.0125)),
    "c5": Workload(
        "c5", dict(seed=5, backbone_bp=100_000_000, snp_rate=1 / 20, indel_rate=1 / 150, indel_len=(1, 50),
                   sv_rate=1 / 20_000, sv_len=(50, 5000)),
        dict(seed=505, n_reads=10_000, length=0, len_mode=2, len_min=100, len_max=300, sub=0.025, ins=0.0125,
             dele=0.0125, mix_frac=0.5, len2=(5000, 50_000), err_scale=(0.2, 1.0)),
        alphabet=N.ALPHABET_MATCH_N),
}


def small_workload(seed: int = 7, backbone_bp: int = 3000, n_reads: int = 8, length: int = 120,
                   error: float = 0.05, n_rate: float = 0.0):
    """A tiny seeded graph + reads for parity tests (oracle finishes in seconds)."""
    ds = Dataset.generate(seed=seed, backbone_bp=backbone_bp, snp_rate=1 / 25, indel_rate=1 / 120,
                          indel_len=(1, 12))
    rs = ReadSet.generate(ds, seed=seed * 31 + 1, n_reads=n_reads, length=length, sub=error / 2, ins=error / 4,
                          dele=error / 4, n_rate=n_rate)
    return ds, rs
