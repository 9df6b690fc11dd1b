"""Access to the golden fixtures written by tests/golden/make_golden.py (result
counts and digests of the reference BPPivot's output)."""
from __future__ import annotations

import json
import os

import numpy as np

from gmkbe_amd.graphio import read_graph_bin

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_G = json.load(open(os.path.join(HERE, "golden.json")))


def graph(name: str):
    return read_graph_bin(os.path.join(HERE, _G["graphs"][name]))


def cases(max_count: int | None = None, min_k: int = 0):
    out = []
    for c in _G["cases"]:
        if max_count is not None and c["count"] > max_count:
            continue
        if c["k"] < min_k:
            continue
        out.append(c)
    return out


def case_id(c) -> str:
    return f"{c['graph']}-k{c['k']}-t{c['theta']}"


def reference_set(c):
    if "set" not in c:
        return None
    return np.fromfile(os.path.join(HERE, c["set"]), dtype="<u4")


def digest_matches(c, dg) -> bool:
    return (int(dg.count) == c["count"] and str(dg.fp_sum) == c["fp_sum"] and str(dg.fp_xor) == c["fp_xor"]
            and str(dg.canon_lo) == c["canon_lo"] and str(dg.canon_hi) == c["canon_hi"])
