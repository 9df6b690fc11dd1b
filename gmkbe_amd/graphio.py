"""Reading and writing the ``.graph.bin`` CSR format.

Layout (all little-endian uint32), as written by the reference converter
(preprocess/graph_conver.cpp:787-809) and read by data/src/tmp_load.h:213-295:

    nEdge (= sum of degrees = 2|E|), N (vertex count), nLeft
    then for i = 0..N-1:  i, deg_i, nbr_0 .. nbr_{deg_i - 1}   (sorted ascending)

Left vertices are 0..nLeft-1, right vertices nLeft..N-1.  File size is
12 + 8N + 4*nEdge bytes.
"""
from __future__ import annotations

import dataclasses
import numpy as np


@dataclasses.dataclass
class Graph:
    n_left: int
    n: int
    off: np.ndarray  # uint32[n + 1]
    adj: np.ndarray  # uint32[off[n]]

    @property
    def n_right(self) -> int:
        return self.n - self.n_left

    @property
    def n_edges(self) -> int:
        return int(self.off[-1]) // 2

    def degrees(self) -> np.ndarray:
        return np.diff(self.off.astype(np.int64)).astype(np.uint32)


def read_graph_bin(path: str) -> Graph:
    raw = np.fromfile(path, dtype="<u4")
    if raw.size < 3:
        raise ValueError(f"{path}: truncated header")
    n_edge, n, n_left = int(raw[0]), int(raw[1]), int(raw[2])
    if n_left > n:
        raise ValueError(f"{path}: bipartite index {n_left} > vertex count {n}")
    if raw.size != 3 + 2 * n + n_edge:
        raise ValueError(f"{path}: size {raw.size * 4} B does not match header (expected {12 + 8 * n + 4 * n_edge})")
    off = np.zeros(n + 1, dtype=np.int64)
    adj = np.empty(n_edge, dtype=np.uint32)
    pos, w = 3, 0
    # records are variable length; walk them with the degree field
    ids = np.empty(n, dtype=np.int64)
    degs = np.empty(n, dtype=np.int64)
    for i in range(n):
        ids[i] = raw[pos]
        d = int(raw[pos + 1])
        degs[i] = d
        adj[w:w + d] = raw[pos + 2:pos + 2 + d]
        w += d
        pos += 2 + d
    if not np.array_equal(ids, np.arange(n)):
        raise ValueError(f"{path}: vertex records are not in id order")
    off[1:] = np.cumsum(degs)
    return Graph(n_left, n, off.astype(np.uint32), adj)


def write_graph_bin(path: str, g: Graph) -> None:
    deg = g.degrees().astype(np.uint32)
    out = np.empty(3 + 2 * g.n + int(g.off[-1]), dtype="<u4")
    out[0], out[1], out[2] = int(g.off[-1]), g.n, g.n_left
    rec_start = 3 + 2 * np.arange(g.n, dtype=np.int64) + g.off[:-1].astype(np.int64)
    out[rec_start] = np.arange(g.n, dtype=np.uint32)
    out[rec_start + 1] = deg
    # neighbour slots: for each vertex, positions rec_start+2 .. +deg
    vid = np.repeat(np.arange(g.n, dtype=np.int64), deg.astype(np.int64))
    within = np.arange(int(g.off[-1]), dtype=np.int64) - g.off[:-1].astype(np.int64)[vid]
    out[rec_start[vid] + 2 + within] = g.adj
    out.tofile(path)


def from_edges(n_left: int, n_right: int, us: np.ndarray, vs: np.ndarray) -> Graph:
    """CSR from edges (u in [0, n_left), v in [0, n_right), right IDs are local);
    duplicate edges are dropped, lists sorted, IDs kept as given (no relabel)."""
    us = np.asarray(us, dtype=np.int64)
    vs = np.asarray(vs, dtype=np.int64)
    key = np.unique(us * n_right + vs)
    us, vs = key // n_right, key % n_right
    n = n_left + n_right
    src = np.concatenate([us, vs + n_left])
    dst = np.concatenate([vs + n_left, us])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off)
    return Graph(n_left, n, off.astype(np.uint32), dst.astype(np.uint32))
