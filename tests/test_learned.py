"""The learned cost model: kernels against a float64 reference, training, and
what the trained model gains over scheduling without it."""
import numpy as np
import pytest

from chsga_amd import learned as L
from chsga_amd.scheduler import fifo_order, plan_order, simulate_ticks


def _params(seed=0):
    rng = np.random.default_rng(seed)
    n = L.N_H1 * L.N_IN + L.N_H1 + L.N_H2 * L.N_H1 + L.N_H2 + L.N_OUT * L.N_H2 + L.N_OUT
    return (rng.standard_normal(n) * 0.3).astype(np.float32)


def test_numpy_reference_gradient_matches_finite_differences():
    """Pins the float64 reference itself: central differences on random
    parameters agree with its analytic gradient (so a dropped ReLU mask, a
    transposed weight or a missing bias term would fail here)."""
    rng = np.random.default_rng(1)
    p = _params(1).astype(np.float64)
    X = rng.standard_normal((17, L.N_IN))
    T = rng.standard_normal((17, L.N_OUT))
    loss, g, _ = L.numpy_forward_backward(p, X, T)
    idx = rng.choice(len(p), size=60, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        num = (L.numpy_forward_backward(p + e, X, T)[0] - L.numpy_forward_backward(p - e, X, T)[0]) / 2e-6
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(g[i])


def test_traceback_bytes_formula():
    # rows 0, k, 2k, ... plus min(k, len-1)+1 segment rows plus the walk buffer
    assert L.traceback_peak_bytes(1001, 10, 500) == 4 * 64 * (3 + 501) + 16 * (2 * 501 + 1024)


def test_lpt_packing_beats_fifo_on_a_lognormal_stream():
    """With exact per-read costs, longest-first packing never needs more ticks
    than arrival order, and on a heavy-tailed stream it needs fewer."""
    rng = np.random.default_rng(4)
    lens = np.clip(np.exp(np.log(3000) + 0.9 * rng.standard_normal(2000)), 500, 30000).astype(int)
    fifo = simulate_ticks(lens, np.arange(len(lens)), 64)
    lpt = simulate_ticks(lens, np.argsort(-lens, kind="stable"), 64)
    lower = int(np.ceil(lens.sum() / 64))
    assert lower <= lpt < fifo


@pytest.mark.gpu
def test_gpu_gradients_match_reference():
    rng = np.random.default_rng(2)
    m = L.CostModel(seed=3)
    p = m.params()
    for n in (1, 5, 300):
        X = rng.standard_normal((n, L.N_IN)).astype(np.float32)
        T = rng.standard_normal((n, L.N_OUT)).astype(np.float32)
        loss, g = m.loss_grad(X, T)
        rl, rg, ry = L.numpy_forward_backward(p, X, T)
        assert abs(loss - rl) <= 1e-4 * max(1.0, abs(rl))
        assert np.allclose(g, rg, rtol=1e-3, atol=1e-5), float(np.abs(g - rg).max())
        assert np.allclose(m.forward(X), ry, rtol=1e-4, atol=1e-5)
    m.close()


@pytest.mark.gpu
def test_gpu_training_fits_a_known_function():
    """y = (log-length + log-size, 2 x log-length): Adam on the GPU drives the
    loss far below that of the initial network."""
    rng = np.random.default_rng(5)
    X = rng.uniform(0.2, 1.0, size=(4000, L.N_IN)).astype(np.float32)
    T = np.stack([X[:, 0] + X[:, 1], 2 * X[:, 0]], axis=1).astype(np.float32)
    m = L.CostModel(seed=1)
    before, _ = m.loss_grad(X, T)
    after = m.fit(X, T, epochs=60, batch=256, lr=3e-3, seed=2)
    assert after < 0.01 * before and after < 1e-3
    m.close()


@pytest.mark.gpu
def test_trained_model_against_the_formula_and_fifo():
    """The committed model (trained on the GPU on 20,000 logged plain runs)
    against the method without it.  A read's plain DP time is its rows times
    a per-row cost set by the graph, so the network learns the per-row cost
    and the plain formula (rows x V) ranks reads exactly as it does on one
    graph: their LPT orders pack the c4 stream in the same number of ticks,
    both below FIFO.  The network gains nothing over its fallback here; its
    held-out error on the logged times is of the same size as a fitted
    two-term formula's (a rows V + b rows), which is the noise of the
    timings."""
    import json
    m = L.load_default()
    if m is None:
        pytest.skip("no trained model committed yet")
    from chsga_amd import WORKLOADS
    w = WORKLOADS["c4"]
    ds = w.dataset()
    rs = w.readset(ds, n_reads=2048)
    fifo = simulate_ticks(rs.lengths(), fifo_order(rs), 64)
    learned = simulate_ticks(rs.lengths(), plan_order(rs, ds, m), 64)
    formula = simulate_ticks(rs.lengths(), plan_order(rs, ds, None), 64)
    print("ticks: fifo", fifo, "formula", formula, "learned", learned, flush=True)
    assert learned == formula < fifo
    rep = json.loads((L.DATA_DIR / "cost_model_report.json").read_text())
    print("held-out median time error: model", rep["test_median_rel_err_time"], "fitted formula",
          rep["formula_median_rel_err_time"], flush=True)
    assert rep["tuples"] >= 20000
    assert rep["test_median_rel_err_time"] < 2 * rep["formula_median_rel_err_time"]
    m.close()
