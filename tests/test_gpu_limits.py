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


@pytest.mark.parametrize("partition,mode", [("global", "batched"), ("seeded", "plain"), ("seeded", "batched")])
def test_result_buffer_overflow_is_reported(partition, mode):
    """A record buffer too small for the results (1000 words for 21,087
    results) ends the call with an overflow error from the expansion (global)
    or the thin checks (seeded); the next call on the same context returns the
    reference's set."""
    import golden_cases as GC
    from gmkbe_amd.engine import GmkError
    c = [c for c in GC.cases() if c["graph"] == "er40x50" and c["k"] == 1 and c["theta"] == 1][0]
    with Engine(0) as e:
        e.set_graph(GC.graph("er40x50"))
        print(f"[gpu] er40x50 k=1 theta=1 {partition} {mode} with a 1000-word record buffer", flush=True)
        with pytest.raises(GmkError, match="result buffer overflow"):
            e.enumerate(k=1, theta=1, mode=mode, partition=partition, capacity_words=1000)
        res = e.enumerate(k=1, theta=1, mode=mode, partition=partition)
        assert GC.digest_matches(c, O.digest(res.words))


def test_asymmetric_csr_is_refused():
    """An edge listed on one side only is refused (the thin checks read side A's
    lists and the partitions side B's, so the two must agree)."""
    import numpy as np
    from gmkbe_amd.engine import GmkError
    from gmkbe_amd.graphio import Graph, from_edges
    g = from_edges(3, 3, np.array([0, 1, 2, 2]), np.array([0, 1, 1, 2]))
    with Engine(0) as e:
        e.set_graph(g)  # symmetric: accepted
        adj = g.adj.copy()
        # left vertex 2 drops right vertex 5 (local 2) from its list; 5 still lists 2
        off = g.off.copy()
        s, t = int(off[2]), int(off[3])
        keep = [x for x in adj[s:t] if x != 5]
        adj2 = np.concatenate([adj[:s], np.array(keep, dtype=np.uint32), adj[t:]])
        off2 = off.copy()
        off2[3:] -= 1
        print("[gpu] asymmetric CSR", flush=True)
        with pytest.raises(GmkError, match="invalid argument"):
            e.set_graph(Graph(g.n_left, g.n, off2, adj2))


def test_limits_around_the_full_count():
    """Limits just below, at and above the full result count of a seeded run
    (launches shrink to the results still allowed, and one may end exactly at
    the limit): exactly min(limit, N) valid, distinct results, the full set
    when the limit is not reached."""
    import numpy as np
    import golden_cases as GC
    c = [c for c in GC.cases() if c["graph"] == "er200x300" and c["k"] == 1 and c["theta"] == 1][0]
    g = GC.graph("er200x300")
    with Engine(0) as e:
        e.set_graph(g)
        full = e.enumerate(k=1, theta=1, mode="batched", partition="seeded")
        n = full.count
        assert GC.digest_matches(c, O.digest(full.words))
        for lim in (n - 70_000, n - 1, n, n + 1):
            print(f"[gpu] er200x300 k=1 seeded limit {lim} (full {n})", flush=True)
            res = e.enumerate(k=1, theta=1, limit=lim, mode="batched", partition="seeded")
            assert res.count == min(lim, n), (lim, res.count)
            assert res.stats["hit_limit"] == (lim <= n)
            f = O.fingerprints(res.words)
            assert np.unique(f).size == res.count
            if lim >= n:
                assert GC.digest_matches(c, O.digest(res.words))
            else:
                assert O.validate(g, 1, 1, O.sample_records(res.words, 5000, seed=lim)) == 0
