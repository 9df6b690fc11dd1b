"""The many-graphs workload as drawn: the first 40 graphs of its seeded draw
(gen.er_batch_specs(2000, SEED_C3): sides log-uniform in [20, 2000], density
log-uniform in [0.01, 0.2], k in {1, 2}), each fully enumerated at the
smallest theta in 1, 2k+1, 2k+3 the reference BPPivot finishes within 60 s
with at most 3M results (tests/golden/c3_slice.json, written by
tests/golden/make_golden.py --c3; 5 of the 40 finish at none of them and are
left out).  One gmk_enumerate_batch call runs the slice in plain, batched and
learned order; every graph's count and order-free digest equal the
reference's."""
import json
import os

import numpy as np
import pytest

from gmkbe_amd import engine, gen
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def slice_():
    spec = json.load(open(os.path.join(HERE, "c3_slice.json")))
    draw = gen.er_batch_specs(spec["draw"], spec["seed"])
    cases = [c for c in spec["cases"] if c["theta"] is not None]
    graphs = []
    for c in cases:
        a, b, p, k, sd = draw[c["index"]]
        assert (a, b, k) == (c["n_left"], c["n_right"], c["k"])
        graphs.append((gen.er(a, b, p, sd), k, c["theta"]))
    return cases, graphs


@pytest.mark.parametrize("mode", ["plain", "batched", "learned"])
def test_c3_slice_matches_reference(slice_, mode):
    cases, graphs = slice_
    assert len(cases) == 35
    with Engine(0) as e:
        if mode == "learned":
            # graph cost model trained on the slice's own per-graph device times
            # in plain mode (a graph run alone)
            print("[gpu] c3 slice: plain run for the cost model", flush=True)
            per, _ = e.enumerate_batch(graphs, mode="plain")
            X = np.stack([engine.graph_features(g, k, t) for g, k, t in graphs])
            y = np.log(np.maximum([p["ms"] for p in per], 1e-3)).astype(np.float32)
            model, _ = e.mlp_train(X, y, epochs=300, lr=5e-3, seed=1)
            e.set_cost_model(model)
        print(f"[gpu] c3 slice: {len(graphs)} graphs, {mode}", flush=True)
        per, tot = e.enumerate_batch(graphs, mode=mode)
    bad = [(c["index"], c["k"], c["theta"], c["count"], p["count"]) for c, p in zip(cases, per)
           if (p["count"], str(p["fp_sum"]), str(p["fp_xor"])) != (c["count"], c["fp_sum"], c["fp_xor"])]
    assert not bad
    assert tot["count"] == sum(c["count"] for c in cases)
