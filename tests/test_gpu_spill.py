"""Stacks that run more than 3/4 full move their shallowest states to the
spill buffer and take them back once drained; the result set is unchanged.
Every stack is capped at 4096 states (GMK_STACK_STATES) and a launch takes at
most 64 states (tuning table), so a small graph's frontier accumulates past
3/4 of the stack: results must still equal the reference's canonical sets, and
the runs must have spilled."""
import os

import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_stacks():
    old = os.environ.get("GMK_STACK_STATES")
    os.environ["GMK_STACK_STATES"] = "4096"
    yield
    if old is None:
        del os.environ["GMK_STACK_STATES"]
    else:
        os.environ["GMK_STACK_STATES"] = old


RUNS = [("er40x50", 1, 1, "global", "batched"), ("er40x50", 2, 5, "global", "plain"),
        ("er200x300", 1, 1, "global", "batched"), ("moreno_crime", 1, 2, "seeded", "batched")]


def test_spilled_searches_are_exact(small_stacks):
    spills = 0
    for name, k, theta, partition, mode in RUNS:
        c = [c for c in GC.cases() if c["graph"] == name and c["k"] == k and c["theta"] == theta][0]
        print(f"[gpu] {name} k={k} theta={theta} {partition} {mode} with 4096-state stacks", flush=True)
        with Engine(0) as e:
            e.set_tuning(dict(e.get_tuning(), batch_states=[64] * 7))
            e.set_graph(GC.graph(name))
            res = e.enumerate(k=k, theta=theta, mode=mode, partition=partition)
        assert GC.digest_matches(c, O.digest(res.words)), (name, k, theta, partition, mode)
        spills += res.stats["spills"]
    print(f"[gpu] spills: {spills}", flush=True)
    assert spills > 0
