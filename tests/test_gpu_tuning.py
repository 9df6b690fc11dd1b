"""Tuning tables change launches, never results: the same batch under
different expansion launch forms (one workgroup per state, persistent grids)
and batch sizes gives identical per-graph counts and fingerprints; values
outside the documented ranges are refused."""
import numpy as np
import pytest

from gmkbe_amd import gen
from gmkbe_amd.engine import Engine, GmkError

pytestmark = pytest.mark.gpu


def _graphs():
    # 12 graphs, 44,850 maximal biplexes in total (theta = 2k + 1), enumerated
    # in full so that whole result sets can be compared
    rng = np.random.default_rng(21)
    out = []
    for i in range(12):
        a, b = int(rng.integers(30, 120)), int(rng.integers(30, 120))
        k = int(rng.integers(1, 3))
        out.append((gen.er(a, b, 0.1, 600 + i), k, 2 * k + 1))
    return out


def test_results_do_not_depend_on_tuning():
    graphs = _graphs()
    sigs = []
    with Engine(0, tuning=None) as e:
        base = e.get_tuning()
        for w, b in [(16, 65536), (0, 65536), (4, 64), (32, 1024)]:
            t = dict(base, waves_per_cu=[w] * 7, batch_states=[b] * 7)
            e.set_tuning(t)
            print(f"[gpu] tuning waves/CU={w} batch={b}", flush=True)
            per, tot = e.enumerate_batch(graphs, mode="batched")
            assert tot["count"] == 44850
            sigs.append([(p["count"], p["fp_sum"], p["fp_xor"]) for p in per])
    assert all(s == sigs[0] for s in sigs[1:])


def test_out_of_range_tuning_is_refused():
    with Engine(0, tuning=None) as e:
        base = e.get_tuning()
        for bad in (dict(base, batch_states=[10] * 7), dict(base, waves_per_cu=[100] * 7),
                    dict(base, dedup_load_permille=950), dict(base, version=99)):
            with pytest.raises(GmkError):
                e.set_tuning(bad)
        assert e.get_tuning() == base
