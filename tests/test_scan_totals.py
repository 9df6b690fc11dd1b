"""Prefix sums of thin family totals (the rank space of a check launch).

Totals saturate at SAT = 2^62 (combination counts of hub tuples exceed 2^64).
A family whose own prefix is finite must keep it even when the running sum
saturates inside it, or the ranks past the previous family decode to offsets
beyond that family's combinations (repeats) and the saturated family is never
enumerated.  The device scan's phases (gmkbe_amd/csrc/thin.h scan_phase_*) run
here on the CPU, thread by thread, against a sequential saturating sum."""
import numpy as np
import pytest

import oracle_lib as O

SAT = 1 << 62


def sequential(totals):
    out, run = [], 0
    for t in totals:
        out.append(run)
        run = min(run + int(t), SAT)
    out.append(run)
    return out


def test_saturating_family_keeps_its_finite_prefix():
    assert [int(x) for x in O.scan_totals_model([10, 20, SAT, 5])] == [0, 10, 30, SAT, SAT]


@pytest.mark.parametrize("n,at", [(3000, 1500), (3000, 1), (5000, 4999), (1024, 512), (100000, 77777)])
def test_saturation_inside_a_segment(n, at):
    """More than 1024 families: each scan thread owns a segment of several, and
    the saturating family falls inside one (not at its start)."""
    rng = np.random.default_rng(n + at)
    totals = rng.integers(0, 1 << 40, size=n, dtype=np.uint64)
    totals[at] = SAT
    got = [int(x) for x in O.scan_totals_model(totals)]
    assert got == sequential(totals)
    assert got[at] < SAT and got[at + 1] == SAT


def test_near_saturation_and_empty():
    assert [int(x) for x in O.scan_totals_model([])] == [0]
    big = [SAT // 3] * 5
    assert [int(x) for x in O.scan_totals_model(big)] == sequential(big)
    rng = np.random.default_rng(7)
    t = rng.integers(0, 1 << 52, size=4097, dtype=np.uint64)
    assert [int(x) for x in O.scan_totals_model(t)] == sequential(t)
