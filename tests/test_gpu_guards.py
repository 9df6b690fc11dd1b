"""Guard tables: the thin checks with the families' guard tables (candidates
from the chosen vertices' lists, free-list misses read from the table, exact
dominance) and without them (GMK_GUARDS=0: the plain candidate lists) give
the reference's canonical sets, and on power-law graphs the same sets as the
CPU model of the kernels."""
import os

import numpy as np
import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd import gen
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu

CASES = [("er12x10", 3, 1), ("er12x10", 2, 2), ("southern_women", 2, 3), ("er40x50", 1, 1),
         ("moreno_crime", 1, 2), ("er200x300", 1, 1)]


@pytest.fixture
def guards():
    """Sets GMK_GUARDS; tables are built for every family (GMK_GUARD_MIN_RANKS=0)
    so the small graphs exercise them."""
    old = {v: os.environ.get(v) for v in ("GMK_GUARDS", "GMK_GUARD_MIN_RANKS")}
    os.environ["GMK_GUARD_MIN_RANKS"] = "0"
    yield lambda v: os.environ.__setitem__("GMK_GUARDS", str(v))
    for v, x in old.items():
        if x is None:
            os.environ.pop(v, None)
        else:
            os.environ[v] = x


@pytest.mark.parametrize("on", [0, 1], ids=["plain", "guards"])
def test_golden_sets_with_and_without_guards(guards, on):
    guards(on)
    with Engine(0) as e:
        for name, k, theta in CASES:
            c = [c for c in GC.cases() if c["graph"] == name and c["k"] == k and c["theta"] == theta][0]
            print(f"[gpu] {name} k={k} theta={theta} seeded, guards {on}", flush=True)
            e.set_graph(GC.graph(name))
            for mode in ("plain", "batched"):
                res = e.enumerate(k=k, theta=theta, mode=mode, partition="seeded")
                assert GC.digest_matches(c, O.digest(res.words)), (name, k, theta, mode, on)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_power_law_thin_sets_match_model(guards, k):
    """Thin results of hub graphs (where tables, families with no candidate
    at all and dominated families occur) equal the CPU model's, with and
    without guard tables."""
    a, b, m = {1: (60, 45, 260), 2: (40, 30, 150), 3: (24, 20, 80)}[k]
    lb = 2 * k + 1
    with Engine(0) as e:
        for seed in (1, 2, 3, 4):
            g = gen.chung_lu(a, b, m, 2.0, 9100 + seed)
            ref, _, _ = O.bppivot(g, k, 1, mode=0)
            want = O.digest(ref)
            e.set_graph(g)
            for on in (0, 1):
                guards(on)
                print(f"[gpu] chung_lu {a}x{b} seed {seed} k={k}, guards {on}", flush=True)
                res = e.enumerate(k=k, theta=1, mode="batched", partition="seeded")
                got = O.digest(res.words)
                assert (got.count, got.canon_lo, got.canon_hi) == (want.count, want.canon_lo, want.canon_hi), (seed, on)


def test_limit_run_with_guards_validates(guards):
    """A limited run on the 50k x 30k power-law graph returns exactly the
    limit, with distinct fingerprints and valid maximal records, with and
    without guard tables."""
    g = gen.workload("cl50kx30k")
    with Engine(0) as e:
        e.set_graph(g)
        for on in (0, 1):
            guards(on)
            print(f"[gpu] cl50kx30k k=1 r=200000, guards {on}", flush=True)
            res = e.enumerate(k=1, theta=1, limit=200_000, mode="batched", partition="seeded")
            assert res.count == 200_000
            fp = O.fingerprints(res.words)
            assert len(np.unique(fp)) == res.count
            assert O.validate(g, 1, 1, O.sample_records(res.words, 2000, seed=on)) == 0
