"""The kernels' combination arithmetic (C(n, c) for c <= 3 saturating at
2^62, and the colex unrank of a combination's rank) against exact integer
arithmetic: math.comb, and the colex rank sum_t C(idx[t], t + 1) == r of
strictly increasing indices below n."""
import math
import random

import oracle_lib as O

SAT = 1 << 62


def test_binom_matches_exact():
    rng = random.Random(5)
    ns = list(range(0, 200)) + [rng.randrange(1 << 32) for _ in range(3000)] + [2 ** 21, 2 ** 21 + 1, 3_037_000_500,
                                                                               (1 << 32) - 1, 1 << 32, 1 << 40]
    for n in ns:
        for c in range(4):
            assert O.model_binom(n, c) == min(math.comb(n, c), SAT), (n, c)


def test_unrank_is_colex_inverse():
    rng = random.Random(7)
    cases = 0
    for _ in range(4000):
        n = rng.choice([rng.randrange(3, 60), rng.randrange(3, 5000), rng.randrange(3, 3_000_000), rng.randrange(3, 1 << 31)])
        for c in (1, 2, 3):
            tot = math.comb(n, c)
            if tot >= SAT:
                continue
            for r in (0, tot - 1, rng.randrange(tot)):
                idx = O.model_unrank(r, c, n)
                assert all(0 <= x < n for x in idx), (n, c, r, idx)
                assert all(a < b for a, b in zip(idx, idx[1:])), (n, c, r, idx)
                assert sum(math.comb(x, t + 1) for t, x in enumerate(idx)) == r, (n, c, r, idx)
                cases += 1
    assert cases > 10_000


def test_unrank_small_enumerates_every_subset_once():
    """Brute force on small n: ranks 0 .. C(n, c) - 1 give every c-subset once."""
    import itertools
    for n in range(3, 13):
        for c in (1, 2, 3):
            got = [tuple(O.model_unrank(r, c, n)) for r in range(math.comb(n, c))]
            assert sorted(got) == sorted(itertools.combinations(range(n), c))
