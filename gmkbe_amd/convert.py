"""Edge list -> ``.graph.bin`` conversion with the reference converter's relabel.

Restates preprocess/graph_conver.cpp:500-817 (``Convert``):

1. every raw ID that occurs in either column is compacted by rank
   (graph_conver.cpp:566-584);
2. IDs seen in the first column become left vertices, IDs seen in the second
   column right vertices, each numbered by compacted rank (graph_conver.cpp:586-613);
   an ID used in both columns yields one left and one right vertex;
3. duplicate edges are dropped per adjacency list (graph_conver.cpp:706-713);
4. each side is relabelled by sorting on (degree_bucket(d'), d', old id) with
   d' = deg - 2 on the left and deg - 1 on the right, computed in uint32, so a
   left vertex of degree 1 wraps to 0xFFFFFFFF and sorts last
   (graph_conver.cpp:715-733; degree_bucket at graph_conver.cpp:443-451);
5. lists are rebuilt under the new IDs and sorted (graph_conver.cpp:746-777).
"""
from __future__ import annotations

import os

import numpy as np

from .graphio import Graph


def degree_bucket(d: np.ndarray) -> np.ndarray:
    """Number of times 32 must be multiplied in before it reaches d, capped at 8
    (graph_conver.cpp:443-451)."""
    d = np.asarray(d, dtype=np.uint64)
    b = np.zeros(d.shape, dtype=np.int64)
    limit = np.uint64(32)
    for _ in range(8):
        more = d > limit
        b += more
        limit = limit * np.uint64(32)
    return b


def parse_edge_text(path: str, threads: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """The raw (u, v) pairs of an edge list, parsed on `threads` host threads
    (0: all cores) by the library's native parser (gmk_parse_edge_list), with
    graph_convert's rules (graph_conver.cpp:505-552)."""
    import ctypes
    from .engine import load_library
    L = load_library()
    up, vp = ctypes.POINTER(ctypes.c_uint32)(), ctypes.POINTER(ctypes.c_uint32)()
    m = ctypes.c_uint64(0)
    rc = L.gmk_parse_edge_list(path.encode(), threads, ctypes.byref(up), ctypes.byref(vp), ctypes.byref(m))
    if rc != 0:
        raise OSError(f"gmk_parse_edge_list({path}): status {rc}")
    try:
        n = int(m.value)
        u = np.ctypeslib.as_array(up, shape=(max(n, 1),))[:n].copy() if n else np.zeros(0, dtype=np.uint32)
        v = np.ctypeslib.as_array(vp, shape=(max(n, 1),))[:n].copy() if n else np.zeros(0, dtype=np.uint32)
    finally:
        L.gmk_free(up)
        L.gmk_free(vp)
    return u, v


def parse_edge_text_py(path: str) -> tuple[np.ndarray, np.ndarray]:
    """The same rules in Python, over the whole file as one range: blanks and
    line ends skipped, '%' / '#' lines are comments, otherwise two numbers are
    read, each skipping any non-digits before it, and the rest of the line is
    dropped (graph_conver.cpp:433-441, 533-549)."""
    data = open(path, "rb").read()
    n, p = len(data), 0
    us, vs = [], []

    def num(p):
        while p < n and not (48 <= data[p] <= 57):
            p += 1
        val = 0
        while p < n and 48 <= data[p] <= 57:
            val = (val * 10 + data[p] - 48) & 0xFFFFFFFF
            p += 1
        return val, p

    while p < n:
        ch = data[p]
        if ch in (10, 13, 32, 9):
            p += 1
            continue
        if ch in (37, 35):  # '%', '#'
            while p < n and data[p] != 10:
                p += 1
            continue
        u, p = num(p)
        v, p = num(p)
        us.append(u)
        vs.append(v)
        while p < n and data[p] != 10:
            p += 1
    return np.asarray(us, dtype=np.uint32), np.asarray(vs, dtype=np.uint32)


def convert_edges(u: np.ndarray, v: np.ndarray) -> Graph:
    u = np.asarray(u, dtype=np.uint32)
    v = np.asarray(v, dtype=np.uint32)
    if u.size == 0:
        raise ValueError("empty edge list")
    ids = np.unique(np.concatenate([u, v]))
    nid = ids.size
    # rank of each raw ID among the IDs in use (a dense table when the IDs are
    # compact, binary search otherwise)
    if int(ids[-1]) < 4 * nid + (1 << 20):
        rank = np.zeros(int(ids[-1]) + 1, dtype=np.int64)
        rank[ids] = np.arange(nid, dtype=np.int64)
        cu = rank[u]
        cv = rank[v]
    else:
        cu = np.searchsorted(ids, u)
        cv = np.searchsorted(ids, v)
    seen_src = np.zeros(nid, dtype=bool)
    seen_dst = np.zeros(nid, dtype=bool)
    seen_src[cu] = True
    seen_dst[cv] = True
    src_rank = np.cumsum(seen_src) - 1
    dst_rank = np.cumsum(seen_dst) - 1
    n_left = int(seen_src.sum())
    n_right = int(seen_dst.sum())
    n = n_left + n_right
    lu = src_rank[cu].astype(np.int64)
    rv = dst_rank[cv].astype(np.int64) + n_left
    # unique undirected pairs (a left-right pair listed twice collapses)
    key = np.unique(lu * n + rv)
    lu, rv = key // n, key % n
    deg = np.zeros(n, dtype=np.int64)
    np.add.at(deg, lu, 1)
    np.add.at(deg, rv, 1)

    def order(start: int, count: int, sub: int) -> np.ndarray:
        old = np.arange(start, start + count, dtype=np.int64)
        dprime = ((deg[old] - sub) & 0xFFFFFFFF).astype(np.uint64)
        return old[np.lexsort((old, dprime, degree_bucket(dprime)))]

    left_nodes = order(0, n_left, 2)
    right_nodes = order(n_left, n_right, 1)
    gmap = np.empty(n, dtype=np.int64)
    gmap[left_nodes] = np.arange(n_left)
    gmap[right_nodes] = n_left + np.arange(n_right)
    a = gmap[lu]
    b = gmap[rv]
    # pairs are distinct, so sorting src * n + dst orders them by (src, dst)
    pair = np.sort(np.concatenate([a * n + b, b * n + a]))
    src, dst = pair // n, pair % n
    off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off)
    return Graph(n_left, n, off.astype(np.uint32), dst.astype(np.uint32))


def convert_file(path: str, threads: int = 0) -> Graph:
    return convert_edges(*parse_edge_text(path, threads))


def write_edge_text(path: str, u: np.ndarray, v: np.ndarray) -> None:
    """``u v`` per line (the converter's input format)."""
    with open(path, "w") as f:
        f.write("\n".join(f"{a} {b}" for a, b in zip(np.asarray(u).tolist(), np.asarray(v).tolist())))
        f.write("\n")


def main(argv=None) -> int:
    """graph_convert -i <edge list> -o <prefix|file.graph.bin> (graph_conver.cpp:453-498, 813-816)."""
    import argparse
    from .graphio import write_graph_bin
    ap = argparse.ArgumentParser(prog="graph_convert")
    ap.add_argument("-i", required=True)
    ap.add_argument("-o", required=True)
    a = ap.parse_args(argv)
    out = a.o if a.o.endswith(".graph.bin") else a.o + ".graph.bin"
    d = os.path.dirname(out)
    if d:
        os.makedirs(d, exist_ok=True)
    g = convert_file(a.i)
    write_graph_bin(out, g)
    print(f"Nodes:   {g.n}\n2|E|  :  {int(g.off[-1])}\nnLeft :  {g.n_left}\nnRight:  {g.n_right}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
