"""The seeded decomposition is exact: thin results with the tuple on the left
(|L| <= 2k), thin results with the tuple on the right (|R| <= 2k, |L| >= 2k+1)
and seed partitions for the rest (both sides >= 2k+1, the reference BPPivot's
seeded regime) together give exactly the reference's result set.  The thin part
runs the kernels' own family/maximality code on the CPU (tests/model)."""
import numpy as np
import pytest

import golden_cases as GC
import oracle_lib as O


@pytest.mark.parametrize("c", GC.cases(max_count=600_000, min_k=1), ids=GC.case_id)
def test_thin_plus_fat_equals_reference(c):
    g = GC.graph(c["graph"])
    k, th = c["k"], c["theta"]
    lb = max(th, 2 * k + 1)
    fat, _, _ = O.bppivot(g, k, lb, mode=1)
    parts = [fat]
    if th <= 2 * k:
        parts.append(O.thin_model(g, k, th, 0, th))
        parts.append(O.thin_model(g, k, th, 1, lb))
    words = np.concatenate(parts)
    assert GC.digest_matches(c, O.digest(words))
    assert O.validate(g, k, th, words) == 0
