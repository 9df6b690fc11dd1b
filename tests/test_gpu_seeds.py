"""Seed partitions' vertex sets built on the device (seed_kernels.hip, one
workgroup per seed) against the host restatement of BPPivot's constructSets
(csrc/seeds.cpp, GMK_DEVICE_SEEDS=0): the same partitions (count, widest
side, search states) and the reference's sets, for single graphs and for a
batch of graphs."""
import json
import os

import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd import gen
from gmkbe_amd.engine import Engine

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES = [c for c in GC.cases() if (c["graph"], c["k"], c["theta"]) in
         {("er200x300", 1, 3), ("er200x300", 2, 5), ("cl600x400", 2, 9), ("cl600x400", 1, 3),
          ("southern_women", 2, 5), ("moreno_crime", 1, 3), ("er40x50", 1, 3)}]


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _run(eng, device, **kw):
    os.environ["GMK_DEVICE_SEEDS"] = "1" if device else "0"
    try:
        return eng.enumerate(**kw)
    finally:
        os.environ.pop("GMK_DEVICE_SEEDS", None)


@pytest.mark.parametrize("c", CASES, ids=GC.case_id)
@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_device_seed_sets_match_host(eng, c, mode):
    eng.set_graph(GC.graph(c["graph"]))
    print(f"[gpu] seed sets {GC.case_id(c)} {mode}", flush=True)
    kw = dict(k=c["k"], theta=c["theta"], partition="seeded", mode=mode)
    dev = _run(eng, True, **kw)
    host = _run(eng, False, **kw)
    for key in ("partitions", "widest_side", "states"):
        assert dev.stats[key] == host.stats[key], key
    assert GC.digest_matches(c, O.digest(dev.words))
    assert dev.stats["launches_kernel"][3] > host.stats["launches_kernel"][3]  # the seed kernel ran


def test_device_seed_sets_in_a_batch(eng):
    """The seeded graphs of the many-graphs slice (theta > 2k, fully
    enumerated, tests/golden/c3_slice.json) in one batched call: with device
    and with host vertex sets, every graph's count and digest are the
    reference's, from the same number of partitions."""
    spec = json.load(open(os.path.join(HERE, "c3_slice.json")))
    draw = gen.er_batch_specs(spec["draw"], spec["seed"])
    cases = [c for c in spec["cases"] if c["theta"] is not None and c["theta"] > 2 * c["k"] and c["count"] < 300_000]
    graphs = []
    for c in cases:
        a, b, p, k, sd = draw[c["index"]]
        graphs.append((gen.er(a, b, p, sd), k, c["theta"]))
    print(f"[gpu] seed sets: {len(graphs)} graphs of the many-graphs slice", flush=True)
    parts = []
    for device in (True, False):
        os.environ["GMK_DEVICE_SEEDS"] = "1" if device else "0"
        try:
            per, tot = eng.enumerate_batch(graphs, mode="batched")
        finally:
            os.environ.pop("GMK_DEVICE_SEEDS", None)
        assert [(p["count"], str(p["fp_sum"]), str(p["fp_xor"])) for p in per] == \
            [(c["count"], c["fp_sum"], c["fp_xor"]) for c in cases]
        parts.append(tot["partitions"])
    assert parts[0] == parts[1]
