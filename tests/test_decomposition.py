"""The seeded decomposition is exact: thin results with the tuple on the left
(|L| <= 2k), thin results with the tuple on the right (|R| <= 2k, |L| >= 2k+1)
and seed partitions for the rest (both sides >= 2k+1, the reference BPPivot's
seeded regime) together give exactly the reference's result set.  The thin part
runs the kernels' own family/maximality code on the CPU (tests/model)."""
import numpy as np
import pytest

import golden_cases as GC
import oracle_lib as O


@pytest.mark.parametrize("c", GC.cases(max_count=600_000, min_k=1), ids=GC.case_id)
def test_thin_plus_fat_equals_reference(c):
    g = GC.graph(c["graph"])
    k, th = c["k"], c["theta"]
    lb = max(th, 2 * k + 1)
    fat, _, _ = O.bppivot(g, k, lb, mode=1)
    parts = [fat]
    if th <= 2 * k:
        parts.append(O.thin_model(g, k, th, 0, th))
        parts.append(O.thin_model(g, k, th, 1, lb))
    words = np.concatenate(parts)
    assert GC.digest_matches(c, O.digest(words))
    assert O.validate(g, k, th, words) == 0


@pytest.mark.parametrize("k", [2, 3])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_dominated_families_skipped_exactly(k, seed):
    """Power-law graphs with hubs: most one-vertex tuples have a vertex that
    misses at most k of their neighbours, so the families that vertex can join
    are dropped before any rank is tested.  The result set stays the reference's."""
    from gmkbe_amd import gen
    g = gen.chung_lu(24, 20, 80, 2.0, 9000 + seed)
    assert O.thin_dominated(g, k, 0) > 0 and O.thin_dominated(g, k, 1) > 0
    lb = 2 * k + 1
    ref, _, _ = O.bppivot(g, k, 1, mode=0)
    fat, _, _ = O.bppivot(g, k, lb, mode=1)
    words = np.concatenate([fat, O.thin_model(g, k, 1, 0, 1), O.thin_model(g, k, 1, 1, lb)])
    a, b = O.digest(words), O.digest(ref)
    assert (a.count, a.canon_lo, a.canon_hi) == (b.count, b.canon_lo, b.canon_hi)
    assert O.validate(g, k, 1, words) == 0


@pytest.mark.parametrize("name", ["er12x10", "southern_women", "er40x50", "moreno_crime"])
def test_neighbour_hash_membership(name):
    """Every hashed vertex's set answers membership like its sorted list, for
    every vertex of the graph (hashes for degrees >= 1 and >= 4)."""
    g = GC.graph(name)
    assert O.hash_check(g, 1) == 0
    assert O.hash_check(g, 4) == 0


@pytest.mark.parametrize("c", [c for c in GC.cases(max_count=600_000, min_k=1) if c["theta"] <= 2 * c["k"]],
                         ids=GC.case_id)
def test_thin_model_with_neighbour_hashes(c):
    """Adjacency tests through neighbour hashes (every vertex with a neighbour,
    then degree >= 3) give the same thin results, in the same order, as sorted
    lists."""
    g = GC.graph(c["graph"])
    k, th = c["k"], c["theta"]
    lb = max(th, 2 * k + 1)
    for side_a, min_b in ((0, th), (1, lb)):
        base = O.thin_model(g, k, th, side_a, min_b)
        for hm in (1, 3):
            assert np.array_equal(O.thin_model(g, k, th, side_a, min_b, hash_min=hm), base), (side_a, hm)



@pytest.mark.parametrize("c", [c for c in GC.cases(max_count=600_000, min_k=1) if c["theta"] <= 2 * c["k"]],
                         ids=GC.case_id)
def test_guard_tables_keep_results(c):
    """With the families' guard tables (candidates from the chosen vertices'
    lists, free-list misses from the table, exact dominance), the thin results
    are the same, in the same order, as with the plain candidate lists."""
    g = GC.graph(c["graph"])
    k, th = c["k"], c["theta"]
    lb = max(th, 2 * k + 1)
    for side_a, min_b in ((0, th), (1, lb)):
        base = O.thin_model(g, k, th, side_a, min_b, guards=False)
        for hm in (0, 3):
            got = O.thin_model(g, k, th, side_a, min_b, hash_min=hm)
            assert np.array_equal(got, base), (side_a, hm)


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_guard_modes_on_power_law_graphs(k, seed):
    """Power-law graphs exercise every guard mode (tables, no candidate at all,
    families dropped as dominated); the thin results match the plain lists',
    and with the fat part they give the reference's set."""
    from gmkbe_amd import gen
    a, b, m = {1: (60, 45, 260), 2: (40, 30, 150), 3: (24, 20, 80)}[k]  # full sets of about 10^4-10^6
    g = gen.chung_lu(a, b, m, 2.0, 9100 + seed)
    lb = 2 * k + 1
    modes = {}
    got = [O.thin_model(g, k, 1, 0, 1, modes=modes)]
    seen = dict(modes)
    got.append(O.thin_model(g, k, 1, 1, lb, modes=modes))
    seen = {m: seen[m] + modes[m] for m in seen}
    assert np.array_equal(got[0], O.thin_model(g, k, 1, 0, 1, guards=False))
    assert np.array_equal(got[1], O.thin_model(g, k, 1, 1, lb, guards=False))
    assert seen["table"] > 0 and seen["dominated"] + seen["allmax"] > 0, seen
    ref, _, _ = O.bppivot(g, k, 1, mode=0)
    fat, _, _ = O.bppivot(g, k, lb, mode=1)
    a, b = O.digest(np.concatenate([fat] + got)), O.digest(ref)
    assert (a.count, a.canon_lo, a.canon_hi) == (b.count, b.canon_lo, b.canon_hi)
