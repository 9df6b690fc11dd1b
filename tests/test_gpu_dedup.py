"""The device dedup table (two independent 64-bit hashes per result, open
addressing, compare-and-swap; the role of the reference's hash tables,
include/hashTable.cuh:857-1078, and iTraversal's result tree,
SGMD22-MBPE/include/iTraversal.h:474-495).

The decomposition emits no repeats, so a test that only shows zero duplicates
cannot tell a working table from one that never finds anything.  Here
GMK_DEDUP_REPLAY=1 makes the library produce results twice (every thin check
launch runs twice; every global root state is pushed twice): without the
table the repeats are emitted, with it exactly the repeats are dropped and the
set is the reference's."""
import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _case(graph, k, theta):
    return [c for c in GC.cases() if c["graph"] == graph and c["k"] == k and c["theta"] == theta][0]


@pytest.mark.parametrize("partition", ["global", "seeded"])
def test_dedup_drops_exactly_the_repeats(monkeypatch, partition):
    c = _case("er40x50", 1, 1)
    g = GC.graph("er40x50")
    monkeypatch.setenv("GMK_DEDUP_REPLAY", "1")
    with Engine(0) as e:
        e.set_graph(g)
        print(f"[gpu] er40x50 replayed, {partition}, no dedup", flush=True)
        rep = e.enumerate(k=1, theta=1, partition=partition, mode="batched")
        assert rep.count > c["count"]
        if partition == "global":
            assert rep.count == 2 * c["count"]
        print(f"[gpu] er40x50 replayed, {partition}, dedup", flush=True)
        res = e.enumerate(k=1, theta=1, partition=partition, mode="batched", dedup=True)
    assert GC.digest_matches(c, O.digest(res.words))
    assert res.stats["dedup_dups"] == rep.count - c["count"]
    assert res.stats["dedup_inserts"] == rep.count
    assert res.stats["dedup_undecided"] == 0


def test_dedup_table_grows_in_unlimited_runs():
    """4,397,076 results through a table that starts at 2^20 slots (load 0.5):
    it doubles, rehashing on the device, and loses nothing."""
    c = _case("er200x300", 1, 1)
    with Engine(0) as e:
        e.set_graph(GC.graph("er200x300"))
        print("[gpu] er200x300 k=1 dedup, unlimited", flush=True)
        res = e.enumerate(k=1, theta=1, partition="global", mode="batched", dedup=True, dedup_load=0.5)
    assert GC.digest_matches(c, O.digest(res.words))
    assert res.stats["dedup_inserts"] == c["count"] and res.stats["dedup_dups"] == 0
