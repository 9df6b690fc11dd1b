"""The converter restatement (gmkbe_amd/convert.py) is byte-exact against the
reference converter's output committed in tests/golden/graphs."""
import os

import numpy as np
import pytest

import golden_cases as GC
from gmkbe_amd import convert, gen, graphio

GOLDEN = {
    "er10x12": lambda: gen.er_edges(10, 12, 0.3, 11),
    "er12x10": lambda: gen.er_edges(12, 10, 0.45, 12),
    "er40x50": lambda: gen.er_edges(40, 50, 0.1, 13),
    "er200x300": lambda: gen.er_edges(200, 300, 0.05, gen.SEED_C1),
    "cl600x400": lambda: gen.chung_lu_edges(600, 400, 3000, 2.1, 17, cap=120),
}


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_converter_byte_exact(name, tmp_path):
    g = convert.convert_edges(*GOLDEN[name]())
    p = tmp_path / "g.graph.bin"
    graphio.write_graph_bin(str(p), g)
    ref = os.path.join(GC.HERE, "graphs", name + ".graph.bin")
    assert open(p, "rb").read() == open(ref, "rb").read()


def test_degree_one_left_vertices_sort_last():
    # left deg-1 vertices wrap to 0xFFFFFFFF under deg - 2 (graph_conver.cpp:721)
    g = convert.convert_edges(np.array([0, 1, 1, 1, 2, 2]), np.array([10, 10, 11, 12, 11, 12]))
    deg = g.degrees()[:g.n_left]
    assert list(deg) == [2, 3, 1]


def test_bucket_edges():
    d = np.array([0, 32, 33, 1024, 1025, 32768, 32769, 0xFFFFFFFF], dtype=np.uint64)
    assert list(convert.degree_bucket(d)) == [0, 0, 1, 1, 2, 2, 3, 6]


def test_roundtrip_read_write(tmp_path):
    g = GC.graph("moreno_crime")
    p = tmp_path / "x.graph.bin"
    graphio.write_graph_bin(str(p), g)
    g2 = graphio.read_graph_bin(str(p))
    assert np.array_equal(g.off, g2.off) and np.array_equal(g.adj, g2.adj) and g.n_left == g2.n_left
