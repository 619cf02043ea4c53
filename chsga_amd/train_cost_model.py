"""Log plain runs on the GPU and train the learned cost model on them.

    python -m chsga_amd.train_cost_model [--tuples N] [--out chsga_amd/data]

1. Seeded graphs of three sizes (the c2/c4 graph and two smaller ones) and
   seeded reads with log-uniform lengths and varying N fractions.
2. Every read is aligned on the plain path; its measured seconds (HIP stream
   synchronised per read) and the peak device bytes of its checkpointed
   traceback at a drawn k form one tuple.  The network learns the time per
   DP row (seconds / rows), so its predictions extrapolate to longer reads.  The bytes of a sample are measured
   by running the traceback and must equal traceback_peak_bytes().
3. The MLP is trained on the GPU (chsga_mlp_train) on log targets and saved;
   its error is reported next to the plain formula's (a * rows * V + b * rows,
   least squares on the same tuples).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def log(*a):
    print(*a, flush=True)


def collect(n_tuples: int, seed: int = 7):
    from chsga_amd import Dataset, Engine, ReadSet
    from chsga_amd import learned as L
    graphs = [
        dict(seed=2, backbone_bp=5_000_000, snp_rate=1 / 30, indel_rate=1 / 300, indel_len=(1, 20)),
        dict(seed=41, backbone_bp=1_000_000, snp_rate=1 / 25, indel_rate=1 / 200, indel_len=(1, 30)),
        dict(seed=42, backbone_bp=200_000, snp_rate=1 / 40, indel_rate=1 / 400, indel_len=(1, 10)),
    ]
    rng = np.random.default_rng(seed)
    X, T, meta = [], [], []
    per_graph = [n_tuples // 2, n_tuples // 4, n_tuples - n_tuples // 2 - n_tuples // 4]
    checked = 0
    for gi, (gp, n) in enumerate(zip(graphs, per_graph)):
        ds = Dataset.generate(**gp)
        info = ds.info()
        g = L.graph_features(info)
        eng = Engine(ds)
        lens = np.exp(rng.uniform(np.log(16), np.log(1024), size=n)).astype(int)
        nfr = rng.choice([0.0, 0.01, 0.05], size=n)
        ks = rng.choice([16, 64, 250, 500, 1000], size=n)
        done = 0
        t0 = time.time()
        for c0 in range(0, n, 256):
            sel = range(c0, min(n, c0 + 256))
            reads = []
            for i in sel:
                r = ReadSet.generate(ds, seed=seed * 1000003 + gi * 100003 + i, n_reads=1, length=int(lens[i]),
                                     sub=0.02, ins=0.01, dele=0.01, n_rate=float(nfr[i]))
                reads.append(r[0])
            rs = ReadSet.from_list(reads)
            _, (tt, _, _) = eng.align_combo(rs, per_read_times=True)
            for j, i in enumerate(sel):
                nf = reads[j].count(b"N") / len(reads[j])
                X.append(L.features([len(reads[j])], [nf], g, int(ks[i]))[0])
                by = L.traceback_peak_bytes(len(reads[j]), info["V"], int(ks[i]))
                T.append([np.log(tt[j]), np.log(by)])
                meta.append((gi, len(reads[j]), info["V"], float(tt[j]), by))
            # a measured sample of the bytes target
            i0 = sel[0]
            tb = eng.align_read_with_traceback(reads[0], k=int(ks[i0]))
            assert eng.last_stats["peak_device_bytes"] == L.traceback_peak_bytes(len(reads[0]), info["V"], int(ks[i0]))
            checked += 1
            done += len(sel)
            log(f"[collect] graph {gi} V={info['V']}: {done}/{n} tuples ({time.time() - t0:.1f}s)")
        eng.close()
    return np.array(X, np.float32), np.array(T, np.float32), np.array(meta, dtype=np.float64), checked


def formula_error(meta: np.ndarray) -> float:
    """Plain cost formula fitted by least squares: t = a*rows*V + b*rows."""
    rows, V, t = meta[:, 1], meta[:, 2], meta[:, 3]
    A = np.stack([rows * V, rows], axis=1)
    coef, *_ = np.linalg.lstsq(A, t, rcond=None)
    pred = A @ coef
    return float(np.median(np.abs(pred - t) / t))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--tuples", type=int, default=8000)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=str(ROOT / "chsga_amd" / "data"))
    ap.add_argument("--from-tuples", default=None,
                    help="train on tuples logged before (a cost_tuples.npz) instead of logging new runs")
    args = ap.parse_args()
    from chsga_amd import learned as L
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    if args.from_tuples:
        with np.load(args.from_tuples, allow_pickle=False) as z:
            X, T, meta = z["X"], z["T"], z["meta"]
        checked = int(json.loads((Path(args.from_tuples).parent / "cost_model_report.json").read_text())
                      .get("bytes_measurements_checked", 0)) if (Path(args.from_tuples).parent /
                                                                  "cost_model_report.json").exists() else 0
    else:
        X, T, meta, checked = collect(args.tuples)
    np.savez(out / "cost_tuples.npz", X=X, T=T, meta=meta)
    n = len(X)
    # the time target per DP row (a read of len chars runs len rows)
    Tt = T.copy()
    Tt[:, 0] = T[:, 0] - np.log(meta[:, 1])
    rng = np.random.default_rng(0)
    perm = rng.permutation(n)
    tr, te = perm[: int(0.8 * n)], perm[int(0.8 * n):]
    m = L.CostModel(seed=11)
    loss = m.fit(X[tr], Tt[tr], epochs=args.epochs, batch=256, lr=args.lr, seed=5)
    pred = m.forward(X[te])
    t_err = float(np.median(np.abs(np.exp(pred[:, 0] - Tt[te, 0]) - 1)))
    b_err = float(np.median(np.abs(np.exp(pred[:, 1] - T[te, 1]) - 1)))
    f_err = formula_error(meta[te])
    m.save(out / "cost_model.npz")
    report = {"tuples": n, "epochs": args.epochs, "lr": args.lr, "train_loss": loss, "test_median_rel_err_time": t_err,
              "test_median_rel_err_bytes": b_err, "formula_median_rel_err_time": f_err,
              "bytes_measurements_checked": checked}
    (out / "cost_model_report.json").write_text(json.dumps(report, indent=1) + "\n")
    log(json.dumps(report))
    return 0


if __name__ == "__main__":
    sys.exit(main())
