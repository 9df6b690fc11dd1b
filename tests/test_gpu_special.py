"""The special graphs that pin k-biplex maximality in closed form, through the
HIP path: K_{a,b} (one result, the whole graph, for every k), the graph with
no edges (C(a, min(k, a)) C(b, min(k, b)) results), the perfect matching M_n
(C(n, 2) results at k = 1), a star with a pendant edge, and a complete block
with isolated vertices on both sides.  Every mode and partition must return
brute force's set, whose closed forms tests/test_oracle.py checks."""
import numpy as np
import pytest

import oracle_lib as O
from gmkbe_amd import graphio
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _complete(a, b):
    us, vs = np.meshgrid(np.arange(a), np.arange(b), indexing="ij")
    return graphio.from_edges(a, b, us.ravel(), vs.ravel())


def _graphs():
    none = np.array([], dtype=np.int64)
    iso_u, iso_v = np.meshgrid(np.arange(3), np.arange(3), indexing="ij")
    return {
        "K4x5": _complete(4, 5),
        "K1x3": _complete(1, 3),
        "empty4x5": graphio.from_edges(4, 5, none, none),
        "empty2x5": graphio.from_edges(2, 5, none, none),
        "M5": graphio.from_edges(5, 5, np.arange(5), np.arange(5)),
        "star": graphio.from_edges(2, 6, np.array([0, 0, 0, 0, 0, 1]), np.array([0, 1, 2, 3, 4, 5])),
        "K3x3+isolated": graphio.from_edges(5, 6, iso_u.ravel(), iso_v.ravel()),
    }


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("name", list(_graphs()))
def test_special_graph_matches_brute(eng, name):
    g = _graphs()[name]
    eng.set_graph(g)
    for k in (0, 1, 2, 3):
        for theta in (1, 2):
            want, _ = O.brute(g, k, theta)
            dw = O.digest(want)
            for partition in ("global", "seeded"):
                for mode in ("plain", "batched"):
                    print(f"[gpu] {name} k={k} theta={theta} {partition} {mode}", flush=True)
                    res = eng.enumerate(k=k, theta=theta, mode=mode, partition=partition)
                    dg = O.digest(res.words)
                    assert (dg.count, dg.canon_lo, dg.canon_hi) == (dw.count, dw.canon_lo, dw.canon_hi), \
                        (name, k, theta, partition, mode, dg.count, dw.count)
