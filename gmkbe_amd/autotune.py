"""Autotuner for the branch-and-bound expansion launch parameters, per size bucket.

The expansion kernel runs one wave per search state, either one workgroup per
state or as a persistent grid of waves per compute unit that takes states from
a device counter.  Which form and how many states go into one launch depend on
the state's bitset width W (1..64 words): wide states hold more LDS per wave
(fewer resident waves) and their searches branch wider.

The size buckets are those of the many-graphs workload (sides 20-63, 64-255,
256-1023, 1024-2000); a bucket's seed partitions have widths W = 1, 2-4, 8-16
and 32-64 words.  Each bucket is calibrated on 24 graphs drawn like that
workload (Erdos-Renyi, density log-uniform in [0.01, 0.2], k in {1, 2},
theta = 2k+1, at most 1000 results per graph) from a seed of its own, with
the largest side in the bucket.  Every candidate setting runs the batch in
batched mode; the fastest is kept for the bucket's widths.  Every candidate
must return the same per-graph counts, and the same fingerprints for graphs
below the result cap: the tuning changes launches, not results.  The thin
checks' launch form (one lane per rank, or waves taking chunks of ranks) is
tuned per k on the workload of that k (50k x 30k Chung-Lu for k = 1,
2M x 1M Chung-Lu for k = 2, 500k x 500k with planted near-bicliques for
k = 3), and the dedup table's load factor on the 50k x 30k workload.

    python -m gmkbe_amd.autotune [--out gmkbe_amd/tune/mi355x.json]

The table is versioned (gmk_tuning.version) and loaded by every Engine.
"""
from __future__ import annotations

import argparse
import datetime
import json
import os
import sys
import time

import numpy as np

from . import gen
from .engine import TUNE_TABLE, TUNING_VERSION, Engine

# (largest side low, high, widths log2 W the bucket's partitions use)
BUCKETS = [(20, 63, (0,)), (64, 255, (1, 2)), (256, 1023, (3, 4)), (1024, 2000, (5, 6))]
WAVES = (0, 16)
BATCH = (16384, 65536)
DEDUP_LOADS = (300, 500, 700)
CHECK_FORMS = {0: "one rank per lane", 1: "chunked"}
# the thin-result workload of each k and its result limit
THIN_WORKLOADS = {1: ("cl50kx30k", 1_000_000), 2: ("cl2Mx1M", 10_000_000), 3: ("planted500k", 10_000_000)}
CAL_SEED = gen.SEED_C3 + 2000


def bucket_graphs(lo: int, hi: int, count: int):
    """The first `count` graphs of the calibration draw whose largest side is in [lo, hi]."""
    out = []
    for a, b, p, k, sd in gen.er_batch_specs(4000, CAL_SEED):
        if lo <= max(a, b) <= hi:
            out.append((gen.er(a, b, p, sd), k, 2 * k + 1))
            if len(out) == count:
                break
    return out


def _log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def tune(graphs_per_bucket: int = 24, cap: int = 1000, reps: int = 1) -> dict:
    eng = Engine(0, tuning=None)
    base = eng.get_tuning()  # the library's built-in table; only the swept fields change
    chosen_w, chosen_b = list(base["waves_per_cu"]), list(base["batch_states"])
    trials = []
    for lo, hi, lanes in BUCKETS:
        graphs = bucket_graphs(lo, hi, graphs_per_bucket)
        ref = None
        best = None
        for w in WAVES:
            for b in BATCH:
                t = dict(base)
                t["waves_per_cu"] = [w] * 7
                t["batch_states"] = [b] * 7
                eng.set_tuning(t)
                eng.enumerate_batch(graphs[:4], limit=cap, mode="batched")  # warm-up
                times = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    per, tot = eng.enumerate_batch(graphs, limit=cap, mode="batched")
                    times.append(time.perf_counter() - t0)
                # counts always match; the sets of graphs below their cap must too
                # (which results a capped graph keeps depends on the work order)
                sig = [(p["count"],) + ((0, 0) if p["hit_cap"] else (p["fp_sum"], p["fp_xor"])) for p in per]
                if ref is None:
                    ref = sig
                if sig != ref:
                    raise RuntimeError(f"bucket {lo}-{hi}: results changed with waves={w} batch={b}")
                s = min(times)
                trials.append({"bucket": [lo, hi], "waves_per_cu": w, "batch_states": b, "s": s,
                               "results": tot["count"], "states": tot["states"]})
                _log(f"[autotune] sides {lo}-{hi} waves/CU={w:2d} batch={b:5d}: {s * 1e3:9.2f} ms "
                     f"({tot['states']} states)")
                if best is None or s < best[0]:
                    best = (s, w, b)
        for lane in lanes:
            chosen_w[lane], chosen_b[lane] = best[1], best[2]
        _log(f"[autotune] sides {lo}-{hi} (W = {', '.join(str(1 << q) for q in lanes)}): waves/CU={best[1]} "
             f"batch={best[2]}")

    # thin checks per k: one lane per rank over the grid, or waves taking chunks
    # of ranks; seeded, batched, on the workload of that k at its result limit
    forms = list(base.get("check_form", [0, 0, 0, 0]))
    for k, (name, limit) in THIN_WORKLOADS.items():
        eng.set_graph(gen.workload(name))
        best_c = None
        for form in CHECK_FORMS:
            t = dict(base)
            t["waves_per_cu"], t["batch_states"] = chosen_w, chosen_b
            t["check_form"] = [form] * 4
            eng.set_tuning(t)
            times = []
            for _ in range(reps + 1):
                t0 = time.perf_counter()
                r = eng.enumerate(k=k, theta=1, limit=limit, mode="batched", partition="seeded", keep=False)
                times.append(time.perf_counter() - t0)
            s = min(times[1:])
            trials.append({"k": k, "workload": name, "check_form": form, "s": s, "results": r.count})
            _log(f"[autotune] thin checks k={k} ({name}) {CHECK_FORMS[form]}: {s * 1e3:.2f} ms")
            if best_c is None or s < best_c[0]:
                best_c = (s, form)
        forms[k] = best_c[1]

    # dedup table load factor on the 50k x 30k workload
    eng.set_graph(gen.workload("cl50kx30k"))
    t = dict(base)
    t["waves_per_cu"], t["batch_states"], t["check_form"] = chosen_w, chosen_b, forms
    eng.set_tuning(t)
    best_load = None
    for load in DEDUP_LOADS:
        times = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            r = eng.enumerate(k=1, theta=1, limit=1_000_000, mode="batched", partition="seeded", keep=False,
                              dedup=True, dedup_load=load / 1000.0)
            times.append(time.perf_counter() - t0)
        s = min(times[1:])
        trials.append({"dedup_load_permille": load, "s": s, "probes": r.stats["dedup_probes"],
                       "inserts": r.stats["dedup_inserts"]})
        _log(f"[autotune] dedup load {load / 1000:.1f}: {s * 1e3:.2f} ms, "
             f"{r.stats['dedup_probes'] / max(1, r.stats['dedup_inserts']):.3f} probes/insert")
        if best_load is None or s < best_load[0]:
            best_load = (s, load)
    eng.close()
    return {
        "version": TUNING_VERSION,
        "target": "gfx950 (MI355X)",
        "generated": datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"),
        "buckets": [{"sides": [lo, hi], "W": [1 << q for q in lanes]} for lo, hi, lanes in BUCKETS],
        "tuning": {"version": TUNING_VERSION, "waves_per_cu": chosen_w, "batch_states": chosen_b,
                   "dedup_load_permille": best_load[1], "check_form": forms},
        "trials": trials,
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m gmkbe_amd.autotune")
    ap.add_argument("--out", default=TUNE_TABLE)
    ap.add_argument("--graphs", type=int, default=24, help="graphs per size bucket")
    a = ap.parse_args(argv)
    table = tune(a.graphs)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps(table["tuning"]))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
