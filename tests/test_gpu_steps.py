"""Branch-and-bound batches planned on the device: the same exact sets
whatever the number of batches run between two host synchronisations
(GMK_STEPS_PER_SYNC) and whether the step sequence is replayed from a
hipGraph or launched kernel by kernel (GMK_GRAPHS).  A graph is captured once
and replayed many times per run, with the host's own copies, spills and
synchronisations between the replays; every run is checked against the
reference's golden digests."""
import os

import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu

CASES = [c for c in GC.cases() if (c["graph"], c["k"], c["theta"]) in
         {("er200x300", 1, 3), ("cl600x400", 2, 9), ("southern_women", 2, 3), ("er40x50", 1, 1)}]


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture
def env():
    saved = {k: os.environ.get(k) for k in ("GMK_STEPS_PER_SYNC", "GMK_GRAPHS", "GMK_STACK_STATES")}
    yield os.environ
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.mark.parametrize("c", CASES, ids=GC.case_id)
@pytest.mark.parametrize("steps,graphs", [(1, "1"), (8, "0"), (8, "1"), (64, "1")])
@pytest.mark.parametrize("partition", ["global", "seeded"])
def test_steps_per_sync_and_graphs_keep_the_exact_set(eng, env, c, steps, graphs, partition):
    env["GMK_STEPS_PER_SYNC"] = str(steps)
    env["GMK_GRAPHS"] = graphs
    eng.set_graph(GC.graph(c["graph"]))
    print(f"[gpu] {GC.case_id(c)} {partition} steps={steps} graphs={graphs}", flush=True)
    res = eng.enumerate(k=c["k"], theta=c["theta"], partition=partition, mode="batched")
    assert GC.digest_matches(c, O.digest(res.words))


def test_small_stacks_spill_between_graph_replays(eng, env):
    """Stacks of 4096 states make the device halt widths that need a spill;
    the host spills them between replays of the captured sequence."""
    c = [c for c in GC.cases() if (c["graph"], c["k"], c["theta"]) == ("er200x300", 1, 3)][0]
    env["GMK_STACK_STATES"] = "4096"
    env["GMK_GRAPHS"] = "1"
    with Engine(0) as e:
        e.set_graph(GC.graph(c["graph"]))
        print("[gpu] er200x300 k=1 theta=3 global, 4096-state stacks", flush=True)
        res = e.enumerate(k=1, theta=3, partition="global", mode="batched")
    assert res.stats["spills"] > 0
    assert GC.digest_matches(c, O.digest(res.words))
