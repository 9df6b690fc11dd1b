"""The gmkbpe command-line driver: result lines, milestone records and the JSON
time record, in the reference's formats (src/main.cu:1709-1715,
include/output.cuh:93-176, include/timer.cuh:98-194)."""
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import golden_cases as GC
import oracle_lib as O
from gmkbe_amd import engine

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(engine.LIB_PATH), "gmkbpe")


def _records_from_lines(out: str):
    words = []
    for ln in out.splitlines():
        m = re.match(r"Result (\d+): X: ([0-9 ]*) \| Y: ([0-9 ]*)$", ln)
        if m:
            L = [int(x) for x in m.group(2).split()]
            R = [int(x) for x in m.group(3).split()]
            words += [len(L), *L, len(R), *R]
    return np.asarray(words, dtype=np.uint32)


def test_cli_full_enumeration_matches_reference(tmp_path):
    c = [c for c in GC.cases(min_k=1) if c["graph"] == "er40x50" and c["k"] == 1 and c["theta"] == 1][0]
    src = os.path.join(GC.HERE, GC._G["graphs"]["er40x50"])
    g = str(tmp_path / "er40x50.graph.bin")
    shutil.copy(src, g)
    p = subprocess.run([CLI, "-f", g, "-k", "1", "-r", "0", "-q", "1"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    words = _records_from_lines(p.stdout)
    assert GC.digest_matches(c, O.digest(words))
    assert "Running Time:" in p.stdout


def test_cli_limit_milestones_and_time_record(tmp_path):
    src = os.path.join(GC.HERE, GC._G["graphs"]["er200x300"])
    g = str(tmp_path / "er200x300.graph.bin")
    shutil.copy(src, g)
    p = subprocess.run([CLI, "-f", g, "-k", "1", "-r", "100000"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    marks = [int(m) for m in re.findall(r"Output (\d+), cost Time: [0-9.e+]+ ms", p.stdout)]
    assert marks == [1, 10, 100, 1000, 10000, 100000]
    rec = json.load(open(tmp_path / "er200x300.graph.gpu1.json"))
    assert set(rec) == {"1"} and set(rec["1"]) == {"100000"} and rec["1"]["100000"] < 5 * 3600 * 1000
    # a run that cannot reach its limit keeps the 5-hour sentinel
    p = subprocess.run([CLI, "-f", g, "-k", "1", "-t", "9", "-r", "100"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    rec = json.load(open(tmp_path / "er200x300.graph.gpu1.json"))
    assert rec["1"]["100"] == 5 * 3600 * 1000 and "100000" in rec["1"]


def test_cli_tuning_table(tmp_path):
    """--tune loads an autotuner table; results do not depend on it, and a
    table the library refuses ends the run with an error."""
    c = [c for c in GC.cases(min_k=1) if c["graph"] == "er40x50" and c["k"] == 1 and c["theta"] == 1][0]
    src = os.path.join(GC.HERE, GC._G["graphs"]["er40x50"])
    g = str(tmp_path / "er40x50.graph.bin")
    shutil.copy(src, g)
    table = json.load(open(engine.TUNE_TABLE))
    alt = dict(table, tuning=dict(table["tuning"], waves_per_cu=[8] * 7, batch_states=[256] * 7))
    bad = dict(table, tuning=dict(table["tuning"], batch_states=[1] * 7))
    for name, t, ok in (("table.json", table, True), ("alt.json", alt, True), ("bad.json", bad, False)):
        path = str(tmp_path / name)
        json.dump(t, open(path, "w"))
        print(f"[gpu] gmkbpe --tune {name}", flush=True)
        p = subprocess.run([CLI, "-f", g, "-k", "1", "-r", "0", "-q", "1", "--tune", path], capture_output=True,
                           text=True, timeout=120)
        if ok:
            assert p.returncode == 0, p.stderr
            assert GC.digest_matches(c, O.digest(_records_from_lines(p.stdout)))
        else:
            assert p.returncode != 0 and "tuning" in p.stderr


def _records_bppivot(out: str, n_left: int):
    """Records from BPPivot-style lines ("Biplex No.n:", "  SL: a, b, ",
    "  SR: c, d, " with right labels local to the right side)."""
    words, L = [], None
    for ln in out.splitlines():
        if ln.startswith("  SL: "):
            L = [int(x) for x in ln[6:].split(",") if x.strip()]
        elif ln.startswith("  SR: "):
            R = [int(x) + n_left for x in ln[6:].split(",") if x.strip()]
            words += [len(L), *L, len(R), *R]
    return np.asarray(words, dtype=np.uint32)


def _records_itraversal(out: str, n_left: int):
    """Records from iTraversal result keys ("x&...y&...", global IDs, left
    side first, each side ascending)."""
    words = []
    for ln in out.splitlines():
        if re.fullmatch(r"(\d+&)+", ln):
            ids = [int(x) for x in ln.split("&") if x]
            L = [x for x in ids if x < n_left]
            R = [x for x in ids if x >= n_left]
            assert ids == sorted(L) + sorted(R)
            words += [len(L), *L, len(R), *R]
    return np.asarray(words, dtype=np.uint32)


@pytest.mark.parametrize("fmt", ["gmkbpe", "bppivot", "itraversal"])
def test_cli_result_formats(tmp_path, fmt):
    """-q 1 result lines in the reference's format and in those of its two CPU
    enumerators each give back the reference's set; the device line names the
    GPU as the reference does."""
    c = [c for c in GC.cases(min_k=1) if c["graph"] == "er40x50" and c["k"] == 1 and c["theta"] == 1][0]
    src = os.path.join(GC.HERE, GC._G["graphs"]["er40x50"])
    g = str(tmp_path / "er40x50.graph.bin")
    shutil.copy(src, g)
    print(f"[gpu] gmkbpe --format {fmt}", flush=True)
    p = subprocess.run([CLI, "-f", g, "-k", "1", "-r", "0", "-q", "1", "--format", fmt], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert re.search(r"^Using GPU: .+, GPU id: 0$", p.stdout, re.M)
    n_left = GC.graph("er40x50").n_left
    parse = {"gmkbpe": lambda o, n: _records_from_lines(o), "bppivot": _records_bppivot,
             "itraversal": _records_itraversal}[fmt]
    words = parse(p.stdout, n_left)
    assert GC.digest_matches(c, O.digest(words))
    if fmt == "bppivot":
        nums = [int(m) for m in re.findall(r"^Biplex No\.(\d+):$", p.stdout, re.M)]
        assert nums == list(range(1, len(nums) + 1)) and len(nums) == c["count"]
