"""Many graphs in one call (gmk_enumerate_batch): per-graph result sets equal
the reference's, in plain, batched and learned order; per-graph caps hold."""
import numpy as np
import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd import engine, gen
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _cases():
    return [c for c in GC.cases(max_count=600_000, min_k=1) if GC.graph(c["graph"]).n_left <= 4096]


@pytest.mark.parametrize("mode", ["plain", "batched", "learned"])
def test_batch_matches_reference(eng, mode):
    cs = _cases()
    graphs = [(GC.graph(c["graph"]), c["k"], c["theta"]) for c in cs]
    print(f"[gpu] batch of {len(graphs)} graphs, {mode}", flush=True)
    per, tot = eng.enumerate_batch(graphs, mode=mode)
    for c, p in zip(cs, per):
        assert p["count"] == c["count"] and str(p["fp_sum"]) == c["fp_sum"] and str(p["fp_xor"]) == c["fp_xor"], \
            GC.case_id(c)
    assert tot["count"] == sum(c["count"] for c in cs)


def test_batch_with_trained_model_and_caps(eng):
    # 40 small ER graphs, 217,305 results in total (largest 87,899), so the
    # uncapped runs stay short
    rng = np.random.default_rng(7)
    graphs = []
    for i in range(40):
        a, b = int(rng.integers(15, 60)), int(rng.integers(15, 60))
        p = float(rng.uniform(0.05, 0.25))
        k = int(rng.integers(1, 3))
        th = 2 * k + 1 + int(rng.integers(0, 2))
        graphs.append((gen.er(a, b, p, 1000 + i), k, th))
    # per-graph device times of a plain run train the model
    print("[gpu] 40-graph batch, plain", flush=True)
    per_plain, _ = eng.enumerate_batch(graphs, mode="plain")
    assert sum(p["count"] for p in per_plain) == 217305
    X = np.stack([engine.graph_features(g, k, t) for g, k, t in graphs])
    y = np.log(np.array([max(p["ms"], 1e-3) for p in per_plain], dtype=np.float32))
    m, _ = eng.mlp_train(X, y, epochs=300, lr=1e-2)
    eng.set_cost_model(m)
    per_l, _ = eng.enumerate_batch(graphs, mode="learned")
    eng.set_cost_model(None)
    for a, b in zip(per_plain, per_l):
        assert (a["count"], a["fp_sum"], a["fp_xor"]) == (b["count"], b["fp_sum"], b["fp_xor"])
    # caps: every graph keeps min(total, cap) results
    cap = 50
    per_c, _ = eng.enumerate_batch(graphs, mode="batched", limit=cap)
    for a, b in zip(per_plain, per_c):
        assert b["count"] == min(a["count"], cap) and b["hit_cap"] == (a["count"] >= cap)


@pytest.mark.parametrize("mode", ["plain", "batched", "learned"])
def test_batch_of_one_global_graph(eng, mode):
    """A batch whose only graph runs as a global partition (theta <= 2k) has no
    seed jobs: it returns that graph's reference set."""
    c = [c for c in GC.cases() if c["graph"] == "er40x50" and c["k"] == 1 and c["theta"] == 1][0]
    print(f"[gpu] batch of one global graph, {mode}", flush=True)
    per, tot = eng.enumerate_batch([(GC.graph("er40x50"), 1, 1)], mode=mode)
    assert (per[0]["count"], str(per[0]["fp_sum"]), str(per[0]["fp_xor"])) == (c["count"], c["fp_sum"], c["fp_xor"])
