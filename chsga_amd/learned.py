"""Learned cost model for scheduling the batched path and traceback.

An FP32 MLP (8 inputs -> 64 -> 64 -> 2, ReLU), trained on the GPU by the
library's own kernels (chsga_mlp_* in include/chsga/abi.h), predicts for a
read on a graph:
  out[0] = log(seconds per DP row of the plain score-only alignment of the
           read; times its rows, the read's seconds)
  out[1] = log(peak device bytes of its checkpointed traceback with interval k)
The time is learned per row, so that a read's predicted time grows with its
length outside the lengths the model was trained on (a network predicting
whole-read seconds from short reads does not extrapolate monotonically to
reads of 30 kbp, and its queue order then packs worse than the formula's).
from features of the read and the graph.  The predictions drive
  * packing: the batched queue takes reads longest-predicted-time first;
  * HBM admission and k: the traceback interval is the largest k whose
    predicted peak bytes fit the budget (fewest segments that fit).
Alignment arithmetic never depends on the model: scores are exact whatever
the order or k.  The plain cost formula rows x V is the fallback when no
model is given and the yardstick of the model's error.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

from . import native as N

DATA_DIR = Path(__file__).resolve().parent / "data"
FEATURES = ("log_len", "log_V", "log_E_per_V", "levels", "log_n_s", "n_frac", "log_k", "log_cells")
N_IN, N_H1, N_H2, N_OUT = 8, 64, 64, 2


def graph_features(info: dict) -> dict:
    return {"V": float(info["V"]), "E": float(info["E"]), "n_s": float(info["n_s"]),
            "levels": float(info["num_layers"])}


def features(lengths, n_fracs, g: dict, k) -> np.ndarray:
    """Feature rows [n, 8] (float32), scaled to O(1)."""
    L = np.asarray(lengths, dtype=np.float64)
    nf = np.asarray(n_fracs, dtype=np.float64)
    kk = np.broadcast_to(np.asarray(k, dtype=np.float64), L.shape)
    X = np.stack([
        np.log(L) / 10.0,
        np.full_like(L, math.log(g["V"]) / 20.0),
        np.full_like(L, math.log(g["E"] / g["V"])),
        np.full_like(L, g["levels"] / 64.0),
        np.full_like(L, math.log(g["n_s"]) / 20.0),
        nf * 10.0,
        np.log(np.minimum(kk, L)) / 10.0,
        np.log(L * g["V"]) / 30.0,
    ], axis=1)
    return X.astype(np.float32)


def read_features(reads, ds_info: dict, k) -> np.ndarray:
    L = reads.lengths()
    nf = np.array([reads[i].count(b"N") / max(1, len(reads[i])) for i in range(len(reads))])
    return features(L, nf, graph_features(ds_info), k)


def traceback_peak_bytes(length: int, V: int, k: int) -> int:
    """Device bytes the checkpointed traceback allocates: the checkpoint rows
    (rows 0, k, 2k, ...), the segment store (min(k, len-1) + 1 rows; rows are
    padded to a multiple of 64 scores) and the walk buffer; equals
    chsga_stats.peak_device_bytes."""
    n_ckpt = (length - 1) // k + 1
    seg_rows = min(k, length - 1) + 1
    walk = 16 * (2 * seg_rows + 1024)
    row_stride = (V + 63) // 64 * 64
    return 4 * row_stride * (n_ckpt + seg_rows) + walk


class CostModel:
    """The MLP on one GPU."""

    def __init__(self, seed: int = 0, device: int = 0, params: np.ndarray | None = None):
        self._lib = N.engine_lib()
        h = C.c_void_p()
        N._check_engine(self._lib.chsga_mlp_create(N_IN, N_H1, N_H2, N_OUT, seed, device, C.byref(h)))
        self._h = h
        self.n_params = int(self._lib.chsga_mlp_num_params(h))
        if params is not None:
            self.set_params(params)

    @staticmethod
    def _fp(a):
        return a.ctypes.data_as(C.POINTER(C.c_float))

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        N._check_engine(self._lib.chsga_mlp_get_params(self._h, self._fp(out), self.n_params))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        N._check_engine(self._lib.chsga_mlp_set_params(self._h, self._fp(p), len(p)))

    def forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), N_OUT), np.float32)
        N._check_engine(self._lib.chsga_mlp_forward(self._h, self._fp(X), len(X), self._fp(Y)))
        return Y

    def loss_grad(self, X: np.ndarray, T: np.ndarray):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        N._check_engine(self._lib.chsga_mlp_loss_grad(self._h, self._fp(X), self._fp(T), len(X), C.byref(loss),
                                                      self._fp(g)))
        return float(loss.value), g

    def fit(self, X: np.ndarray, T: np.ndarray, epochs: int = 200, batch: int = 256, lr: float = 2e-3,
            seed: int = 0) -> float:
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = C.c_float()
        N._check_engine(self._lib.chsga_mlp_train(self._h, self._fp(X), self._fp(T), len(X), epochs, batch, lr,
                                                  seed, C.byref(loss)))
        return float(loss.value)

    # --- scheduling -------------------------------------------------------
    def predict_time(self, reads, ds, k: int = 500) -> np.ndarray:
        return np.exp(self.forward(read_features(reads, ds.info(), k))[:, 0]) * reads.lengths().astype(np.float64)

    def predict_bytes(self, reads, ds, k: int) -> np.ndarray:
        return np.exp(self.forward(read_features(reads, ds.info(), k))[:, 1])

    def choose_k(self, read: bytes, ds, budget_bytes: float, candidates=(2000, 1000, 500, 250, 125, 64, 32, 16)) -> int:
        """Largest checkpoint interval whose predicted peak bytes fit the budget."""
        from .api import ReadSet
        rs = ReadSet.from_list([read])
        info = ds.info()
        for k in candidates:
            if np.exp(self.forward(read_features(rs, info, k))[0, 1]) <= budget_bytes:
                return k
        return candidates[-1]

    def save(self, path) -> None:
        np.savez(path, params=self.params(), features=np.array(FEATURES))

    @classmethod
    def load(cls, path, device: int = 0) -> "CostModel":
        with np.load(path, allow_pickle=False) as z:
            return cls(device=device, params=z["params"])

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.chsga_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()


def numpy_forward_backward(params: np.ndarray, X: np.ndarray, T: np.ndarray):
    """float64 reference of the MLP's loss and gradient (used to check the
    GPU kernels; the tests pin it by finite differences)."""
    p = params.astype(np.float64)
    X = X.astype(np.float64)
    T = T.astype(np.float64)
    o = 0
    W1 = p[o:o + N_H1 * N_IN].reshape(N_H1, N_IN); o += N_H1 * N_IN
    b1 = p[o:o + N_H1]; o += N_H1
    W2 = p[o:o + N_H2 * N_H1].reshape(N_H2, N_H1); o += N_H2 * N_H1
    b2 = p[o:o + N_H2]; o += N_H2
    W3 = p[o:o + N_OUT * N_H2].reshape(N_OUT, N_H2); o += N_OUT * N_H2
    b3 = p[o:o + N_OUT]
    Z1 = X @ W1.T + b1
    A1 = np.maximum(Z1, 0)
    Z2 = A1 @ W2.T + b2
    A2 = np.maximum(Z2, 0)
    Y = A2 @ W3.T + b3
    n = X.shape[0] * N_OUT
    loss = float(((Y - T) ** 2).sum() / n)
    dY = 2 * (Y - T) / n
    dW3 = dY.T @ A2
    db3 = dY.sum(0)
    dZ2 = (dY @ W3) * (Z2 > 0)
    dW2 = dZ2.T @ A1
    db2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (Z1 > 0)
    dW1 = dZ1.T @ X
    db1 = dZ1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return loss, g, Y


def default_model_path() -> Path:
    return DATA_DIR / "cost_model.npz"


def load_default(device: int = 0) -> "CostModel | None":
    p = default_model_path()
    return CostModel.load(p, device) if p.exists() else None
