"""The multi-core CPU port of the thin-result method (oracle/cpu_port.cpp, the
benchmark's multi-core CPU line) finds exactly the one-vertex thin results
the CPU model of the kernels finds, with any thread count, and stops at a
limit."""
import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd import gen


@pytest.mark.parametrize("k", [1, 2])
@pytest.mark.parametrize("threads", [1, 4])
def test_port_counts_the_one_vertex_results(k, threads):
    """Against the one-vertex records of the CPU model of the kernels."""
    g = GC.graph("er10x12")
    want = sum(1 for L, R in O.records(O.thin_model(g, k, 1, 0, 1)) if len(L) == 1)
    assert want > 0
    got, sec, ranks = O.cpu_thin_port(g, k, threads=threads)
    assert got == want and ranks >= got


def test_port_stops_at_the_limit():
    g = gen.chung_lu(3000, 2000, 20000, 2.1, 7)
    n, sec, ranks = O.cpu_thin_port(g, 1, limit=5000, threads=4)
    assert n == 5000
