"""The thin checks' launch forms and deferred pass give the same result sets.

The first check kernel (either launch form) tests a combination's first
candidates in one lane; one still open after GMK_CHECK_BUDGET of them goes to
a second kernel in which a whole wave tests the rest.  Budgets 0 (every open combination deferred, the wave tests
all its candidates), 1 and 3 push most combinations of the golden graphs
through that second kernel, 64 (the default) few: with every launch form,
results must still equal the reference's canonical sets."""
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
def budget():
    old = os.environ.get("GMK_CHECK_BUDGET")

    def set_budget(v):
        os.environ["GMK_CHECK_BUDGET"] = str(v)

    yield set_budget
    if old is None:
        os.environ.pop("GMK_CHECK_BUDGET", None)
    else:
        os.environ["GMK_CHECK_BUDGET"] = old


@pytest.mark.parametrize("form", [0, 1], ids=["lane", "chunked"])
@pytest.mark.parametrize("b", [0, 1, 3, 64])
def test_deferred_checks_are_exact(budget, b, form):
    budget(b)
    with Engine(0) as e:
        e.set_tuning(dict(e.get_tuning(), check_form=[form] * 4))
        for name, k, theta in CASES:
            c = [c for c in GC.cases() if c["graph"] == name and c["k"] == k and c["theta"] == theta][0]
            print(f"[gpu] {name} k={k} theta={theta} seeded, check form {form}, budget {b}", flush=True)
            e.set_graph(GC.graph(name))
            for mode in ("plain", "batched"):
                res = e.enumerate(k=k, theta=theta, mode=mode, partition="seeded")
                assert GC.digest_matches(c, O.digest(res.words)), (name, k, theta, mode, b)


@pytest.fixture
def wave1():
    old = os.environ.get("GMK_WAVE1_TUPLES")
    yield lambda v: os.environ.__setitem__("GMK_WAVE1_TUPLES", str(v))
    if old is None:
        os.environ.pop("GMK_WAVE1_TUPLES", None)
    else:
        os.environ["GMK_WAVE1_TUPLES"] = old


@pytest.mark.parametrize("limit", [0, 1 << 30], ids=["lane", "wave"])
def test_one_vertex_tuple_build_forms_are_exact(wave1, limit):
    """One-vertex tuples are built a lane each (GMK_WAVE1_TUPLES=0) or a wave
    each (every launch below the limit): the same canonical sets."""
    wave1(limit)
    with Engine(0) as e:
        for name, k, theta in CASES:
            c = [c for c in GC.cases() if c["graph"] == name and c["k"] == k and c["theta"] == theta][0]
            print(f"[gpu] {name} k={k} theta={theta} seeded, wave build below {limit} tuples", flush=True)
            e.set_graph(GC.graph(name))
            res = e.enumerate(k=k, theta=theta, mode="batched", partition="seeded")
            assert GC.digest_matches(c, O.digest(res.words)), (name, k, theta, limit)

