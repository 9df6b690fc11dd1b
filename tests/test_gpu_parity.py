"""GPU parity: the HIP enumerator against the reference BPPivot's golden sets,
brute force and the oracle, through the C ABI."""
import os

import numpy as np
import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd import gen, graphio
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _check_set(g, k, theta, res, expect_words=None, expect_case=None):
    dg = O.digest(res.words)
    assert dg.count == res.count
    assert (dg.fp_sum, dg.fp_xor) == (res.stats["fp_sum"], res.stats["fp_xor"])
    if expect_case is not None:
        assert GC.digest_matches(expect_case, dg), (dg.count, expect_case["count"])
    if expect_words is not None:
        de = O.digest(expect_words)
        assert (dg.count, dg.canon_lo, dg.canon_hi) == (de.count, de.canon_lo, de.canon_hi)


@pytest.mark.parametrize("c", GC.cases(min_k=1), ids=GC.case_id)
def test_global_partition_matches_reference(eng, c):
    g = GC.graph(c["graph"])
    print(f"[gpu] golden {GC.case_id(c)} n={g.n}", flush=True)
    eng.set_graph(g)
    res = eng.enumerate(k=c["k"], theta=c["theta"], partition="global")
    _check_set(g, c["k"], c["theta"], res, expect_case=c)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_global_partition_matches_brute(eng, seed, k):
    rng = np.random.default_rng(100 + seed)
    a, b = int(rng.integers(4, 10)), int(rng.integers(4, 10))
    u, v = gen.er_edges(a, b, float(rng.uniform(0.25, 0.7)), 100 + seed)
    g = graphio.from_edges(a, b, u, v - a)
    eng.set_graph(g)
    for theta in (1, 2, 3):
        print(f"[gpu] brute seed={seed} k={k} theta={theta} {a}x{b}", flush=True)
        wb, _ = O.brute(g, k, theta)
        res = eng.enumerate(k=k, theta=theta, partition="global")
        _check_set(g, k, theta, res, expect_words=wb)


def test_limit_counts_and_validity(eng):
    g = GC.graph("er200x300")
    eng.set_graph(g)
    for lim in (1, 10, 1000, 123457):
        res = eng.enumerate(k=1, theta=1, limit=lim, partition="global")
        assert res.count == lim and res.stats["hit_limit"]
        assert O.validate(g, 1, 1, res.words) == 0


@pytest.mark.parametrize("mode", ["plain", "batched"])
@pytest.mark.parametrize("c", GC.cases(min_k=1), ids=GC.case_id)
def test_seeded_partition_matches_reference(eng, c, mode):
    g = GC.graph(c["graph"])
    if c["k"] > 3 and c["theta"] <= 2 * c["k"]:
        pytest.skip("thin results need k <= 3")
    if mode == "plain" and c["count"] > 600_000:
        pytest.skip("plain mode on the large cases is covered by the batched run")
    print(f"[gpu] seeded {mode} {GC.case_id(c)}", flush=True)
    eng.set_graph(g)
    res = eng.enumerate(k=c["k"], theta=c["theta"], partition="seeded", mode=mode)
    _check_set(g, c["k"], c["theta"], res, expect_case=c)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_seeded_partition_matches_brute(eng, seed, k):
    rng = np.random.default_rng(200 + seed)
    a, b = int(rng.integers(4, 10)), int(rng.integers(4, 10))
    u, v = gen.er_edges(a, b, float(rng.uniform(0.25, 0.7)), 200 + seed)
    g = graphio.from_edges(a, b, u, v - a)
    eng.set_graph(g)
    for theta in (1, 2, 3, 4):
        print(f"[gpu] seeded brute seed={seed} k={k} theta={theta} {a}x{b}", flush=True)
        wb, _ = O.brute(g, k, theta)
        res = eng.enumerate(k=k, theta=theta, partition="seeded", mode="batched")
        _check_set(g, k, theta, res, expect_words=wb)


@pytest.mark.parametrize("partition", ["global", "seeded"])
def test_dedup_table_finds_no_repeats(eng, partition):
    c = [c for c in GC.cases(min_k=1) if c["graph"] == "er40x50" and c["k"] == 1 and c["theta"] == 1][0]
    g = GC.graph(c["graph"])
    eng.set_graph(g)
    res = eng.enumerate(k=1, theta=1, partition=partition, mode="batched", dedup=True, dedup_load=0.7)
    _check_set(g, 1, 1, res, expect_case=c)
    assert res.stats["dedup_inserts"] == c["count"] and res.stats["dedup_dups"] == 0
    assert res.stats["dedup_probes"] >= res.stats["dedup_inserts"]


def _core_dead(g, alpha):
    """Vertices outside the (alpha, alpha)-core: repeatedly remove vertices of
    degree < alpha (numpy restatement of BPPivot's coreReduction,
    biplex.cpp:123-174)."""
    deg = g.degrees().astype(np.int64)
    dead = deg < alpha
    changed = True
    while changed:
        live_deg = np.zeros(g.n, dtype=np.int64)
        src = np.repeat(np.arange(g.n), g.degrees().astype(np.int64))
        keep = ~dead[g.adj.astype(np.int64)] & ~dead[src]
        np.add.at(live_deg, src[keep], 1)
        new_dead = dead | (live_deg < alpha)
        changed = bool((new_dead != dead).any())
        dead = new_dead
    return dead


def _butterfly(g, dead, q, k):
    """The device butterfly reduction restated on the host: BPPivot's rule
    (biplex.cpp:176-225) in passes over one side at a time, every vertex of
    the side decided against the removals of the previous pass, two rounds
    over both sides (the second skipped when the first removed nothing).
    Returns (vertices removed, edges removed between surviving vertices)."""
    nbr = [[int(w) for w in g.adj[g.off[v]:g.off[v + 1]] if not dead[int(w)]] if not dead[v] else []
           for v in range(g.n)]
    deg = [len(x) for x in nbr]
    rm = [d == 0 for d in deg]
    cut = set()
    for _ in range(2):
        any_change = False
        for lo, hi in ((0, g.n_left), (g.n_left, g.n)):
            old, new = list(rm), list(rm)
            for u in range(lo, hi):
                if old[u]:
                    continue
                cn = {}
                for v in nbr[u]:
                    if not old[v]:
                        for w in nbr[v]:
                            if not old[w]:
                                cn[w] = cn.get(w, 0) + 1
                support = 0
                for v in nbr[u]:
                    if old[v]:
                        continue
                    c = sum(1 for w in nbr[v] if not old[w] and cn.get(w, 0) >= q - 2 * k)
                    if c >= q - k:
                        support += 1
                    else:
                        cut.add((min(u, v), max(u, v)))
                        any_change = True
                if support < q - k:
                    new[u] = True
                    any_change = True
                    for v in nbr[u]:
                        deg[v] -= 1
                        if deg[v] == q - k - 1:
                            new[v] = True
            rm = new
        if not any_change:
            break
    removed = sum(1 for v in range(g.n) if rm[v] and nbr[v])
    edges = sum(1 for u, v in cut if not rm[u] and not rm[v])
    return removed, edges


@pytest.mark.parametrize("name,k,theta", [("cl600x400", 1, 3), ("cl600x400", 2, 9), ("moreno_crime", 1, 3),
                                          ("er200x300", 2, 5)])
def test_core_reduction_removes_the_core_complement(eng, name, k, theta):
    """Seeded runs peel the (lb - k)-core on the device before seeding: the
    count of removed vertices equals a host peel's, and the set is still the
    reference's (butterfly reduction off)."""
    c = [c for c in GC.cases() if c["graph"] == name and c["k"] == k and c["theta"] == theta][0]
    g = GC.graph(name)
    eng.set_graph(g)
    print(f"[gpu] core reduction {name} k={k} theta={theta}", flush=True)
    os.environ["GMK_BUTTERFLY"] = "0"
    try:
        res = eng.enumerate(k=k, theta=theta, partition="seeded", mode="batched")
    finally:
        os.environ.pop("GMK_BUTTERFLY", None)
    lb = max(theta, 2 * k + 1)
    assert res.stats["reduce_removed"] == int(_core_dead(g, lb - k).sum())
    assert res.stats["ms_reduce"] > 0
    _check_set(g, k, theta, res, expect_case=c)


@pytest.mark.parametrize("name,k,theta", [("cl600x400", 1, 3), ("cl600x400", 2, 9), ("moreno_crime", 1, 3),
                                          ("er200x300", 2, 5), ("er200x300", 1, 3), ("southern_women", 1, 3)])
def test_butterfly_reduction_matches_host_restatement(eng, name, k, theta):
    """After the core, seeded runs apply the butterfly reduction on the device:
    the vertices and edges it removes equal the host restatement's, and the
    set is still the reference's."""
    c = [c for c in GC.cases() if c["graph"] == name and c["k"] == k and c["theta"] == theta][0]
    g = GC.graph(name)
    eng.set_graph(g)
    print(f"[gpu] butterfly reduction {name} k={k} theta={theta}", flush=True)
    res = eng.enumerate(k=k, theta=theta, partition="seeded", mode="batched")
    lb = max(theta, 2 * k + 1)
    dead = _core_dead(g, lb - k)
    removed, edges = _butterfly(g, dead, lb, k)
    assert res.stats["reduce_removed"] == int(dead.sum()) + removed
    assert res.stats["reduce_edges"] == edges
    _check_set(g, k, theta, res, expect_case=c)
