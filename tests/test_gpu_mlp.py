"""The cost model's HIP kernels against the float64 CPU oracle
(oracle/mlp_oracle.py): forward pass, gradients of every parameter tensor by
central finite differences, and that training lowers the loss."""
import os
import sys

import numpy as np
import pytest

from gmkbe_amd.engine import Engine, Mlp

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 8)).astype(np.float32) * 2 + 1
    t = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 7]).astype(np.float32)
    return X, t


def test_forward_matches_oracle(eng):
    X, t = _data(200, 1)
    m, _ = eng.mlp_train(X, t, epochs=50, lr=1e-2, seed=3)
    got = eng.mlp_predict(m, X)
    ref = mlp_oracle.forward(m.arrays(), X)
    np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-4)


def test_gradients_match_finite_differences(eng):
    X, t = _data(48, 2)
    m, _ = eng.mlp_train(X, t, epochs=20, lr=1e-2, seed=5)  # a non-trivial point with active units
    g, loss = eng.mlp_grad(m, X, t)
    p = m.arrays()
    assert abs(loss - mlp_oracle.loss(p, X, t)) <= 1e-4 * max(1.0, loss)
    ga = g.arrays()
    rng = np.random.default_rng(0)
    checked = 0
    for name in ("w1", "b1", "w2", "b2", "w3", "b3"):
        shape = np.shape(p[name])
        idxs = [()] if shape == () else [tuple(int(rng.integers(0, d)) for d in shape) for _ in range(12)]
        for ix in idxs:
            fd = mlp_oracle.finite_difference(p, X, t, name, ix)
            an = float(ga[name][ix]) if shape else float(ga[name])
            assert abs(an - fd) <= 2e-4 + 2e-3 * abs(fd), (name, ix, an, fd)
            checked += 1
    assert checked >= 61


def test_training_lowers_loss(eng):
    X, t = _data(2000, 4)
    var = float(np.var(t))
    _, loss = eng.mlp_train(X, t, epochs=400, lr=1e-2, seed=1)
    assert loss < 0.2 * var


def _spearman(a, b):
    ra = np.argsort(np.argsort(a)).astype(np.float64)
    rb = np.argsort(np.argsort(b)).astype(np.float64)
    return float(np.corrcoef(ra, rb)[0, 1])


def test_trained_model_ranks_cost_better_than_size_heuristic(eng):
    """What the trained network gains over the method without it: on graphs it
    did not see, its predicted costs order the graphs closer to their measured
    device times than the size heuristic (|E| * n) that orders work without it.
    Targets are the per-graph device times the enumerator logs."""
    from gmkbe_amd import engine, gen
    rng = np.random.default_rng(11)
    graphs = []
    for i in range(100):
        a = int(round(np.exp(rng.uniform(np.log(20), np.log(240)))))
        b = int(round(np.exp(rng.uniform(np.log(20), np.log(240)))))
        p = float(np.exp(rng.uniform(np.log(0.03), np.log(0.25))))
        k = int(rng.integers(1, 3))
        graphs.append((gen.er(a, b, p, 5000 + i), k, 2 * k + 1))
    print("[gpu] 100 graphs, each run alone to log its device time", flush=True)
    ms = []
    for gkt in graphs:
        per, _ = eng.enumerate_batch([gkt], limit=1000, mode="batched")
        ms.append(max(per[0]["ms"], 1e-3))
    y = np.log(np.array(ms, dtype=np.float32))
    X = np.stack([engine.graph_features(g, k, t) for g, k, t in graphs])
    model, _ = eng.mlp_train(X[:70], y[:70], epochs=600, lr=5e-3, seed=1)
    pred = eng.mlp_predict(model, X[70:])
    heur = np.array([np.log(g.n_edges * g.n + 1.0) for g, _, _ in graphs[70:]])
    s_model, s_heur = _spearman(pred, y[70:]), _spearman(heur, y[70:])
    print(f"[gpu] held-out rank correlation: model {s_model:.3f}, size heuristic {s_heur:.3f}", flush=True)
    assert s_model > s_heur and s_model > 0.6
