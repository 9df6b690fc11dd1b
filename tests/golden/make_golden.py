"""Regenerates the golden fixtures in tests/golden/ from the reference's own CPU
tools, compiled from their sources by oracle/ref.mk into oracle/_ref/.

Run in the build container only (it reads /root/reference and runs the
reference binaries there):

    make -f oracle/ref.mk && python tests/golden/make_golden.py

Outputs (data only):
  graphs/<name>.graph.bin   written by the reference converter (graph_convert)
                            from the edge list this script writes
  golden.json               per (graph, k, theta): result count and digests of
                            the reference BPPivot's printed result set
  c3_slice.json             (--c3) per graph of the first 40 of the many-graphs
                            workload's draw: theta and the reference's digest
  wide.json                 (--wide) graphs with seed partitions of 4,500 to
                            40,000 vertices a side (gen.hub): the reference's
                            count and digests; for k = 0 also those of the
                            maximal bicliques by intersection closure
  sets/<case>.u32           the full record stream for small cases
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from gmkbe_amd import convert, gen, graphio  # noqa: E402
import oracle_lib  # noqa: E402

REF = "/root/reference"
BIN = os.path.join(ROOT, "oracle", "_ref")


def ref_convert(name: str, u: np.ndarray, v: np.ndarray) -> str:
    os.makedirs(os.path.join(HERE, "graphs"), exist_ok=True)
    out = os.path.join(HERE, "graphs", name)
    with tempfile.TemporaryDirectory() as td:
        txt = os.path.join(td, name + ".txt")
        convert.write_edge_text(txt, u, v)
        subprocess.run([os.path.join(BIN, "graph_convert"), "-i", txt, "-o", out], check=True,
                       capture_output=True)
    return out + ".graph.bin"


def strip_header_edges(path: str) -> tuple[np.ndarray, np.ndarray]:
    """BPPivot's text fixtures start with a count line; the converter would read
    it as an edge, so it is dropped."""
    lines = open(path).read().split("\n")[1:]
    us, vs = [], []
    for ln in lines:
        p = ln.split()
        if len(p) >= 2:
            us.append(int(p[0])); vs.append(int(p[1]))
    return np.asarray(us, dtype=np.uint32), np.asarray(vs, dtype=np.uint32)


def example_graph_edges(path: str) -> tuple[np.ndarray, np.ndarray]:
    """SGMD22-MBPE/Example.graph: header `N nLeft m`, then one adjacency line per
    vertex with global IDs.  Edges are taken from the left vertices' lines."""
    rows = [list(map(int, ln.split())) for ln in open(path) if ln.strip()]
    n, n_left = rows[0][0], rows[0][1]
    us, vs = [], []
    for r in rows[1:]:
        if r[0] < n_left:
            for w in r[1:]:
                us.append(r[0]); vs.append(w)
    return np.asarray(us, dtype=np.uint32), np.asarray(vs, dtype=np.uint32)


def _unlimit_stack():
    import resource
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_bppivot(graph_path: str, k: int, theta: int, timeout: int = 600, limit: int = 0) -> np.ndarray:
    """Runs the reference BPPivot with -o 1 and returns its result set as a
    record stream with global IDs (right labels are id - nLeft in its output:
    load_bipartite.hpp:58).  The binary is the direct child (no shell), so a
    timeout ends it; limit > 0 passes -n limit."""
    g = graphio.read_graph_bin(graph_path)
    with tempfile.TemporaryDirectory() as td:
        gp = os.path.join(td, "g.graph.bin")
        with open(graph_path, "rb") as f, open(gp, "wb") as o:
            o.write(f.read())
        out = os.path.join(td, "out.txt")
        cmd = [os.path.join(BIN, "bppivot"), "-d", gp, "-q", str(theta), "-k", str(k), "-o", "1"]
        if limit:
            cmd += ["-n", str(limit)]
        with open(out, "w") as fo:
            subprocess.run(cmd, stdout=fo, check=True, timeout=timeout, preexec_fn=_unlimit_stack)
        Ls, Rs = [], []
        with open(out) as f:
            for ln in f:
                s = ln.strip()
                if s.startswith("SL:"):
                    Ls.append(s[3:])
                elif s.startswith("SR:"):
                    Rs.append(s[3:])
        parts: list[np.ndarray] = []
        for ls, rs in zip(Ls, Rs):
            L = np.sort(np.array(ls.replace(",", " ").split(), dtype=np.uint32))
            R = np.sort(np.array(rs.replace(",", " ").split(), dtype=np.uint32)) + np.uint32(g.n_left)
            parts.append(np.concatenate([[len(L)], L, [len(R)], R]).astype(np.uint32))
    return np.concatenate(parts) if parts else np.zeros(0, dtype=np.uint32)


def main() -> None:
    graphs = {}
    graphs["southern_women"] = ref_convert(
        "southern_women", *strip_header_edges(f"{REF}/baselines/MaximalBiPlex/datas/opsahl_southernwomen.txt"))
    graphs["moreno_crime"] = ref_convert(
        "moreno_crime", *strip_header_edges(f"{REF}/baselines/MaximalBiPlex/datas/moreno_crime.txt"))
    graphs["example"] = ref_convert("example", *example_graph_edges(f"{REF}/baselines/SGMD22-MBPE/Example.graph"))
    graphs["er10x12"] = ref_convert("er10x12", *gen.er_edges(10, 12, 0.3, 11))
    graphs["er12x10"] = ref_convert("er12x10", *gen.er_edges(12, 10, 0.45, 12))
    graphs["er40x50"] = ref_convert("er40x50", *gen.er_edges(40, 50, 0.1, 13))
    graphs["er200x300"] = ref_convert("er200x300", *gen.er_edges(200, 300, 0.05, gen.SEED_C1))
    graphs["cl600x400"] = ref_convert("cl600x400", *gen.chung_lu_edges(600, 400, 3000, 2.1, 17, cap=120))

    cases = [
        ("example", 1, 1), ("example", 2, 1), ("example", 1, 2),
        ("er10x12", 1, 1), ("er10x12", 2, 1), ("er10x12", 1, 3), ("er12x10", 1, 1), ("er12x10", 2, 2),
        ("er12x10", 0, 1), ("er12x10", 3, 1),
        ("southern_women", 1, 1), ("southern_women", 2, 3), ("southern_women", 1, 3), ("southern_women", 2, 5),
        ("moreno_crime", 1, 3), ("moreno_crime", 2, 5), ("moreno_crime", 1, 2),
        ("er40x50", 1, 1), ("er40x50", 1, 3), ("er40x50", 2, 5),
        ("cl600x400", 1, 3), ("cl600x400", 2, 9),
        ("er200x300", 1, 1), ("er200x300", 1, 3), ("er200x300", 2, 5),
    ]
    os.makedirs(os.path.join(HERE, "sets"), exist_ok=True)
    gpath = os.path.join(HERE, "golden.json")
    old = json.load(open(gpath)) if os.path.exists(gpath) else {"cases": []}
    have = {(c["graph"], c["k"], c["theta"]): c for c in old["cases"]}
    out = {"graphs": {k: os.path.relpath(v, HERE) for k, v in graphs.items()}, "cases": []}
    for name, k, theta in cases:
        if (name, k, theta) in have and "--force" not in sys.argv:
            out["cases"].append(have[(name, k, theta)])
            continue
        words = run_bppivot(graphs[name], k, theta)
        dg = oracle_lib.digest(words)
        rec = {"graph": name, "k": k, "theta": theta, "count": int(dg.count), "fp_sum": str(dg.fp_sum),
               "fp_xor": str(dg.fp_xor), "canon_lo": str(dg.canon_lo), "canon_hi": str(dg.canon_hi)}
        if dg.count <= 20000:
            fn = f"sets/{name}_k{k}_t{theta}.u32"
            words.astype("<u4").tofile(os.path.join(HERE, fn))
            rec["set"] = fn
        out["cases"].append(rec)
        print(name, k, theta, dg.count, flush=True)
        with open(gpath, "w") as f:
            json.dump(out, f, indent=1)


C3_SLICE = 40          # the first graphs of the many-graphs workload's draw
C3_MAX_COUNT = 3_000_000
C3_TIMEOUT = 60


def c3_case(i: int):
    """Digest of graph i of the many-graphs workload (gen.er_batch_specs with
    its committed seed) at the smallest theta in 1, 2k+1, 2k+3 at which the
    reference BPPivot finishes within C3_TIMEOUT seconds with at most
    C3_MAX_COUNT results."""
    a, b, p, k, sd = gen.er_batch_specs(2000, gen.SEED_C3)[i]
    g = gen.er(a, b, p, sd)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "g.graph.bin")
        graphio.write_graph_bin(path, g)
        for theta in (1, 2 * k + 1, 2 * k + 3):
            try:
                words = run_bppivot(path, k, theta, timeout=C3_TIMEOUT, limit=C3_MAX_COUNT + 1)
            except subprocess.TimeoutExpired:
                continue
            dg = oracle_lib.digest(words)
            if dg.count > C3_MAX_COUNT:
                continue
            return {"index": i, "n_left": a, "n_right": b, "p": p, "k": k, "theta": theta, "count": int(dg.count),
                    "fp_sum": str(dg.fp_sum), "fp_xor": str(dg.fp_xor)}
    return {"index": i, "n_left": a, "n_right": b, "p": p, "k": k, "theta": None}


def main_c3() -> None:
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(8) as ex:
        cases = list(ex.map(c3_case, range(C3_SLICE)))
    for c in cases:
        print(c["index"], c["k"], c["theta"], c.get("count"), flush=True)
    with open(os.path.join(HERE, "c3_slice.json"), "w") as f:
        json.dump({"seed": gen.SEED_C3, "draw": 2000, "max_count": C3_MAX_COUNT, "timeout_s": C3_TIMEOUT,
                   "cases": cases}, f, indent=1)


# Graphs with wide seed partitions (gen.hub: a hub adjacent to every left
# vertex): (name, n_u, n_v, blocks, block side, seed, k, theta).  Their left
# seed partitions hold the whole left side: 4,500 to 40,000 vertices.
WIDE = [
    ("hub4500x1000", 4500, 1000, 3, 6, 3, 1, 3),
    ("hub9000x1000", 9000, 1000, 3, 6, 6, 1, 3),
    ("hub5000x2000", 5000, 2000, 5, 6, 1, 0, 2),
    ("hub9000x2000", 9000, 2000, 5, 6, 2, 0, 2),
    ("hub20000x3000", 20000, 3000, 5, 6, 4, 0, 2),
    ("hub40000x3000", 40000, 3000, 5, 6, 5, 0, 2),
]


def closure_digest(g, theta: int) -> dict:
    """Maximal bicliques by intersection closure (oracle_lib.closure_bicliques):
    the k = 0 check of these graphs, where the reference BPPivot's output
    misses some maximal bicliques (each closure record is checked maximal by
    the oracle's validator)."""
    words = oracle_lib.closure_bicliques(g, theta)
    if oracle_lib.validate(g, 0, theta, words) != 0:
        raise SystemExit("closure records fail the validator")
    dg = oracle_lib.digest(words)
    return {"count": int(dg.count), "fp_sum": str(dg.fp_sum), "fp_xor": str(dg.fp_xor),
            "canon_lo": str(dg.canon_lo), "canon_hi": str(dg.canon_hi)}


def main_wide() -> None:
    wpath = os.path.join(HERE, "wide.json")
    have = {c["name"]: c for c in json.load(open(wpath))["cases"]} if os.path.exists(wpath) else {}
    cases = []
    for name, a, b, blocks, side, seed, k, theta in WIDE:
        if name in have and "--force" not in sys.argv:
            c = have[name]
            if k == 0 and "closure" not in c:
                c["closure"] = closure_digest(gen.hub(a, b, blocks, side, seed), theta)
            cases.append(c)
            continue
        g = gen.hub(a, b, blocks, side, seed)
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "g.graph.bin")
            graphio.write_graph_bin(path, g)
            words = run_bppivot(path, k, theta, timeout=600)
        dg = oracle_lib.digest(words)
        c = {"name": name, "n_u": a, "n_v": b, "blocks": blocks, "side": side, "seed": seed, "k": k,
             "theta": theta, "edges": int(g.n_edges), "count": int(dg.count), "fp_sum": str(dg.fp_sum),
             "fp_xor": str(dg.fp_xor), "canon_lo": str(dg.canon_lo), "canon_hi": str(dg.canon_hi)}
        if k == 0:
            c["closure"] = closure_digest(g, theta)
        cases.append(c)
        print(name, k, theta, dg.count, flush=True)
    with open(os.path.join(HERE, "wide.json"), "w") as f:
        json.dump({"cases": cases}, f, indent=1)


if __name__ == "__main__":
    if "--c3" in sys.argv:
        main_c3()
    elif "--wide" in sys.argv:
        main_wide()
    else:
        main()
