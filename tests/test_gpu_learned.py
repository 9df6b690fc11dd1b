"""The learned schedule of single-graph runs (GMK_MODE_LEARNED with an item
model): the per-item log it trains on, the fused feature and scoring kernel,
exact result sets whatever the order, and what the trained model gains over
the heuristic (degree) order on the 50k x 30k graph."""
import numpy as np
import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd import gen, schedule
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _features_numpy(g, k, first, count):
    """The item features (gmk_item_features) restated in numpy: degree moments
    over at most 16 neighbours spread evenly over N(u)."""
    deg = g.degrees().astype(np.int64)
    nb = g.n - g.n_left if first < g.n_left else g.n_left
    out = np.zeros((count, 8), dtype=np.float64)
    for i in range(count):
        u = first + i
        o, d = int(g.off[u]), int(deg[u])
        ns = min(d, 16)
        idx = np.arange(ns) if ns == d else (np.arange(ns, dtype=np.int64) * d) // ns
        dv = deg[g.adj[o + idx].astype(np.int64)] if ns else np.zeros(0, dtype=np.int64)
        lg = np.log1p(dv.astype(np.float64))
        mean = lg.mean() if ns else 0.0
        var = (lg * lg).mean() - mean * mean if ns else 0.0
        out[i] = [np.log1p(d), np.log1p(nb - d), mean, lg.max() if ns else 0.0, np.sqrt(max(var, 0.0)),
                  np.log1p((dv - 1).sum() * d / ns if ns else 0.0), np.log1p(dv.min() if ns else 0.0), k]
    return out


def test_item_features_match_numpy(eng):
    g = gen.chung_lu(3000, 2000, 20000, 2.0, 77)
    eng.set_graph(g)
    for first, count in ((0, g.n_left), (g.n_left, g.n - g.n_left)):
        print(f"[gpu] item features {first}+{count}", flush=True)
        got = eng.item_features(first, count, 2)
        np.testing.assert_allclose(got, _features_numpy(g, 2, first, count), rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("k", [1, 2])
def test_probe_counts_every_one_vertex_result(eng, k):
    """With a rank budget above every family's size, the probe's per-item
    results are the one-vertex thin results of each vertex, as the CPU model
    of the thin kernels (tests/model) enumerates them."""
    g = GC.graph("er40x50")
    eng.set_graph(g)
    print(f"[gpu] probe er40x50 k={k}", flush=True)
    res, ms = eng.probe_items(0, g.n_left, k, ranks=1 << 40)
    words = O.thin_model(g, k, 1, 0, 1)
    want = np.zeros(g.n_left, dtype=np.int64)
    for L, R in O.records(words):
        if len(L) == 1:
            want[L[0]] += 1
    assert np.array_equal(res.astype(np.int64), want)
    assert (ms[res > 0] > 0).all()


def _train(eng, g, k, seed=1):
    eng.set_graph(g)
    X, y, _ = schedule.probe_samples(eng, k, 0, g.n_left)
    model, _ = schedule.train_item_model(eng, [(X, y)], epochs=200, seed=seed)
    return model


@pytest.mark.parametrize("c", [c for c in GC.cases(max_count=600_000, min_k=1) if c["theta"] == 1], ids=GC.case_id)
def test_learned_order_keeps_the_exact_set(eng, c):
    g = GC.graph(c["graph"])
    model = _train(eng, g, c["k"])
    eng.set_item_model(model)
    try:
        print(f"[gpu] learned {GC.case_id(c)}", flush=True)
        res = eng.enumerate(k=c["k"], theta=1, mode="learned", partition="seeded")
    finally:
        eng.set_item_model(None)
    assert GC.digest_matches(c, O.digest(res.words))


def test_item_scores_match_the_model(eng):
    """The scheduling kernel's fused features + forward pass gives the scores
    gmk_mlp_predict gives on the same features."""
    g = gen.chung_lu(3000, 2000, 20000, 2.0, 78)
    eng.set_graph(g)
    X = eng.item_features(0, g.n_left, 2)
    model, _ = eng.mlp_train(X, np.log1p(X[:, 0] * X[:, 2]).astype(np.float32), epochs=100, lr=1e-2, seed=4)
    eng.set_item_model(model)
    try:
        got = eng.item_predict(0, g.n_left, 2)
    finally:
        eng.set_item_model(None)
    np.testing.assert_allclose(got, eng.mlp_predict(model, X), rtol=1e-4, atol=1e-4)


def test_learned_run_starts_with_the_best_scored_items(eng):
    """The first launch of a learned run checks the first 64 items of the
    order: every result of a run stopped by a small limit belongs to an item
    scored at least as high as the 64th best."""
    g = gen.chung_lu(400, 300, 3000, 2.0, 5)
    eng.set_graph(g)
    X = eng.item_features(0, g.n_left, 1)
    model, _ = eng.mlp_train(X, (X[:, 0] - 0.3 * X[:, 2]).astype(np.float32), epochs=300, lr=1e-2, seed=2)
    eng.set_item_model(model)
    try:
        score = eng.item_predict(0, g.n_left, 1)
        print("[gpu] learned first items", flush=True)
        res = eng.enumerate(k=1, theta=1, limit=50, mode="learned", partition="seeded")
    finally:
        eng.set_item_model(None)
    cut = np.sort(score)[::-1][63]
    heads = [L[0] for L, R in O.records(res.words)]
    assert len(heads) == 50 and all(len(L) == 1 for L, R in O.records(res.words))
    assert all(score[u] >= cut for u in heads)


def test_learned_model_gains_over_the_heuristic_order(eng):
    """50k x 30k graph, k = 1, r = 10^6: with the trained item model the run
    checks far fewer combinations to reach the limit than in the heuristic
    order (measured 1.26M against 4.2M), and every result is valid."""
    g = gen.workload("cl50kx30k")
    model = _train(eng, g, 1)
    print("[gpu] cl50kx30k heuristic order", flush=True)
    base = eng.enumerate(k=1, theta=1, limit=1_000_000, mode="learned", partition="seeded", fetch=False)
    eng.set_item_model(model)
    try:
        print("[gpu] cl50kx30k learned order", flush=True)
        res = eng.enumerate(k=1, theta=1, limit=1_000_000, mode="learned", partition="seeded")
    finally:
        eng.set_item_model(None)
    assert res.count == base.count == 1_000_000
    assert res.stats["intersections"] < 0.6 * base.stats["intersections"]
    dg = O.digest(res.words)
    assert dg.count == res.count and (dg.fp_sum, dg.fp_xor) == (res.stats["fp_sum"], res.stats["fp_xor"])
    assert O.validate(g, 1, 1, res.words) == 0
