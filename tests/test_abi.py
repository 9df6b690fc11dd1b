"""The C-ABI library loads and exports every entry point include/gmk.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from gmkbe_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "gmk.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gmk_[a-z_]+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(engine.LIB_PATH)
    names = declared_symbols()
    assert "gmk_enumerate" in names and len(names) >= 7
    for name in names:
        assert hasattr(lib, name), name


def test_engine_fails_loudly_without_gpu_or_library(monkeypatch):
    import pytest
    monkeypatch.setattr(engine, "LIB_PATH", "/nonexistent/libgmk.so")
    monkeypatch.setattr(engine, "_lib", None)
    with pytest.raises(engine.GmkError):
        engine.Engine(0)


def test_params_and_stats_layout_match_header(tmp_path):
    """The ctypes mirrors have the C structs' sizes and field offsets, as the C
    compiler lays out include/gmk.h."""
    import ctypes
    import subprocess
    structs = {"gmk_params": engine.Params, "gmk_stats": engine.Stats, "gmk_tuning": engine.Tuning,
               "gmk_graph_desc": engine.GraphDesc, "gmk_graph_stats": engine.GraphStats, "gmk_mlp": engine.Mlp}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{os.path.join(ROOT, "include", "gmk.h")}"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(ln.split()[:2]): int(ln.split()[2]) for ln in out if ln}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_committed_tuning_table_is_valid():
    """The table every Engine loads holds the current version and values the
    library accepts (gmk_set_tuning's ranges), and survives the ctypes round
    trip unchanged."""
    import json
    with open(engine.TUNE_TABLE) as f:
        t = json.load(f)["tuning"]
    assert t["version"] == engine.TUNING_VERSION
    assert len(t["waves_per_cu"]) == 7 and all(0 <= w <= 64 for w in t["waves_per_cu"])
    assert len(t["batch_states"]) == 7 and all(64 <= b <= 65536 for b in t["batch_states"])
    assert 100 <= t["dedup_load_permille"] <= 900
    assert len(t["check_form"]) == 4 and all(f in (0, 1) for f in t["check_form"])
    assert engine.Tuning.from_dict(t).as_dict() == t
