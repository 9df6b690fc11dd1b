"""Limited runs on the measured workloads: exactly r results, every one a
distinct maximal k-biplex of the graph (oracle validator), digest consistent."""
import pytest

import oracle_lib as O
from gmkbe_amd import gen
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    return gen.workload("cl50kx30k")


@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_cl50kx30k_limit_valid(c2, mode):
    print(f"[gpu] cl50kx30k k=1 r=1000000 {mode}", flush=True)
    with Engine(0) as e:
        e.set_graph(c2)
        res = e.enumerate(k=1, theta=1, limit=1_000_000, mode=mode, partition="seeded")
    assert res.count == 1_000_000 and res.stats["hit_limit"]
    dg = O.digest(res.words)
    assert dg.count == res.count and (dg.fp_sum, dg.fp_xor) == (res.stats["fp_sum"], res.stats["fp_xor"])
    assert O.validate(c2, 1, 1, res.words) == 0


@pytest.mark.parametrize("k,theta,limit", [(2, 1, 200_000), (1, 2, 200_000), (2, 3, 100_000)])
def test_cl50kx30k_other_k_valid(c2, k, theta, limit):
    print(f"[gpu] cl50kx30k k={k} theta={theta} r={limit}", flush=True)
    with Engine(0) as e:
        e.set_graph(c2)
        res = e.enumerate(k=k, theta=theta, limit=limit, mode="batched", partition="seeded")
    assert res.count == limit
    assert O.validate(c2, k, theta, res.words) == 0
