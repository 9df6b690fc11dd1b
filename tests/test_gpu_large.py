"""The large single-graph workloads at full size, to their limits: the
2M x 1M Chung-Lu graph at k = 2 and the 500k x 500k graph with planted
near-bicliques at k = 3, r = 10,000,000 each, in plain, batched and learned
mode (the item model trained as the benchmark trains it).
Every run returns exactly r results, all distinct (per-record fingerprints),
with the digest the device accumulated, and a spread sample of 2000 records
passes the oracle validator (k-biplex, maximal in the whole graph, sides >= 1).
Full validation of 10^7 records is out of reach on the CPU here: records hold
vertices of degree up to 10^6, and the validator scans every neighbour list."""
import numpy as np
import pytest

import oracle_lib as O
from gmkbe_amd import gen
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _check(g, k, limit, mode):
    with Engine(0) as e:
        if mode == "learned":
            # the item model as the benchmark trains it: the 50k x 30k graph's
            # items and this graph's first 5%
            from gmkbe_amd import schedule
            samples = []
            for gt, kt, frac in ((gen.workload("cl50kx30k"), 1, 1.0), (g, k, 0.05)):
                e.set_graph(gt)
                X, y, _ = schedule.probe_samples(e, kt, 0, int(gt.n_left * frac))
                samples.append((X, y))
            model, _ = schedule.train_item_model(e, samples)
            e.set_item_model(model)
        e.set_graph(g)
        res = e.enumerate(k=k, theta=1, limit=limit, mode=mode, partition="seeded")
    assert res.count == limit and res.stats["hit_limit"]
    f = O.fingerprints(res.words)
    assert f.size == limit and np.unique(f).size == limit
    assert int(f.sum(dtype=np.uint64)) == res.stats["fp_sum"]
    assert int(np.bitwise_xor.reduce(f)) == res.stats["fp_xor"]
    assert O.validate(g, k, 1, O.sample_records(res.words, 2000, seed=7)) == 0


@pytest.fixture(scope="module")
def c4():
    print("[gpu] generating the 2M x 1M Chung-Lu graph", flush=True)
    return gen.workload("cl2Mx1M")


@pytest.fixture(scope="module")
def c5():
    print("[gpu] generating the 500k x 500k graph with planted near-bicliques", flush=True)
    return gen.workload("planted500k")


@pytest.mark.parametrize("mode", ["plain", "batched", "learned"])
def test_cl2Mx1M_k2_limit(c4, mode):
    print(f"[gpu] cl2Mx1M k=2 r=10000000 {mode}", flush=True)
    _check(c4, 2, 10_000_000, mode)


@pytest.mark.parametrize("mode", ["plain", "batched", "learned"])
def test_planted500k_k3_limit(c5, mode):
    print(f"[gpu] planted500k k=3 r=10000000 {mode}", flush=True)
    _check(c5, 3, 10_000_000, mode)
