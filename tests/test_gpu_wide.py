"""Partitions wider than 4096 vertices a side (64 bitset words).  On the
graphs of gen.hub a hub is adjacent to every left vertex, so with k = 1 every
left seed's partition holds the whole left side, and the whole graph is one
partition of 5,000 to 40,000 left vertices.  Widths up to
128 words keep a state's working set in LDS; wider ones (256, 512, 1024
words) keep it in an HBM slice per persistent wave.  Each run's count and
digests equal the reference BPPivot's on the same graph (k = 1), or those of
the maximal bicliques by intersection closure (k = 0), from
tests/golden/wide.json (written by tests/golden/make_golden.py --wide)."""
import json
import os

import pytest

import oracle_lib as O
from gmkbe_amd import gen
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {c["name"]: c for c in json.load(open(os.path.join(HERE, "wide.json")))["cases"]}


def _graph(c):
    g = gen.hub(c["n_u"], c["n_v"], c["blocks"], c["side"], c["seed"])
    assert g.n_edges == c["edges"]
    return g


def _check(want, res):
    dg = O.digest(res.words)
    assert (dg.count, str(dg.fp_sum), str(dg.fp_xor), str(dg.canon_lo), str(dg.canon_hi)) == (
        want["count"], want["fp_sum"], want["fp_xor"], want["canon_lo"], want["canon_hi"])
    assert (str(res.stats["fp_sum"]), str(res.stats["fp_xor"])) == (want["fp_sum"], want["fp_xor"])


# (case, mode, the least widest side the run must show).  With k = 1 and
# theta = 3 every left vertex (adjacent to the hub) stays in every left seed's
# partition; with k = 0 the partitions shrink to the seed's neighbourhood, so
# the k = 0 graphs run as one global partition.
# Plain mode (one partition at a time) is left out here: with k = 1 these
# searches are chains of include/exclude branches as deep as the partition
# (thousands of steps per seed, each a launch and a synchronisation), which
# batched mode advances for all seeds at once.
SEEDED = [("hub4500x1000", "batched", 4097), ("hub9000x1000", "batched", 8193)]


@pytest.mark.parametrize("name,mode,widest", SEEDED, ids=[f"{n}-{m}" for n, m, _ in SEEDED])
def test_wide_seed_partitions_match_reference(name, mode, widest):
    c = CASES[name]
    g = _graph(c)
    print(f"[gpu] wide {name} k={c['k']} theta={c['theta']} seeded {mode}", flush=True)
    with Engine(0) as eng:  # a context of its own: its wide state stacks are freed at the end
        eng.set_graph(g)
        res = eng.enumerate(k=c["k"], theta=c["theta"], mode=mode, partition="seeded")
    assert res.stats["widest_side"] >= widest
    _check(c, res)


@pytest.mark.parametrize("name", ["hub5000x2000", "hub9000x2000", "hub20000x3000", "hub40000x3000"])
def test_wide_global_partition_matches_closure(name):
    """The whole graph as one partition: 128, 256, 512 and 1024 words on the
    left (the last three with their working sets in HBM), k = 0.  The check is
    the set of maximal bicliques by intersection closure: on these graphs the
    reference BPPivot's k = 0 output misses 8 to 19 maximal bicliques (each
    closure record passes the oracle's maximality validator, and the closure
    equals brute force on small graphs, tests/test_oracle.py)."""
    c = CASES[name]
    g = _graph(c)
    print(f"[gpu] wide {name} k={c['k']} theta={c['theta']} global", flush=True)
    with Engine(0) as eng:
        eng.set_graph(g)
        res = eng.enumerate(k=c["k"], theta=c["theta"], mode="batched", partition="global")
    assert res.stats["widest_side"] == c["n_u"]
    _check(c["closure"], res)
