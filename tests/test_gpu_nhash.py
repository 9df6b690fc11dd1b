"""Neighbour hashes in the thin checks give the same result sets.

Uploading a graph builds an open-addressing set of neighbours for every vertex
of degree >= GMK_NHASH_MIN (0: none); the thin checks' adjacency tests use it
instead of binary search.  With hashes for every vertex that has a neighbour
(1), for degrees >= 4, and with none (0), results must equal the reference's
canonical sets."""
import os

import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu

# (graph, k, theta): cases with results of at most 2k vertices on one side
CASES = [("er12x10", 3, 1), ("er12x10", 2, 2), ("southern_women", 2, 3), ("er40x50", 1, 1),
         ("moreno_crime", 1, 2), ("er200x300", 1, 1)]


@pytest.fixture
def hash_min():
    old = os.environ.get("GMK_NHASH_MIN")

    def set_min(v):
        os.environ["GMK_NHASH_MIN"] = str(v)

    yield set_min
    if old is None:
        os.environ.pop("GMK_NHASH_MIN", None)
    else:
        os.environ["GMK_NHASH_MIN"] = old


@pytest.mark.parametrize("hm", [0, 1, 4])
def test_hashed_adjacency_is_exact(hash_min, hm):
    hash_min(hm)
    with Engine(0) as e:
        for name, k, theta in CASES:
            c = [c for c in GC.cases() if c["graph"] == name and c["k"] == k and c["theta"] == theta][0]
            print(f"[gpu] {name} k={k} theta={theta} seeded, neighbour hashes from degree {hm}", flush=True)
            e.set_graph(GC.graph(name))
            for mode in ("plain", "batched"):
                res = e.enumerate(k=k, theta=theta, mode=mode, partition="seeded")
                assert GC.digest_matches(c, O.digest(res.words)), (name, k, theta, mode, hm)
