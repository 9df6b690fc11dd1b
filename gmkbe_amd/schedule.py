"""Training the item cost model of the learned schedule (GMK_MODE_LEARNED
single-graph runs).

The method logs its own per-item yield and device time with gmk_probe_items:
for each one-vertex work item (a vertex u with N(u) and up to k vertices it
misses) the first `ranks` combinations of each of its families are checked,
giving the maximal results found and the lane time spent.  The model learns
log((results + 1) / (lane ms + ITEM_MS0)) from the item's features
(gmk_item_features), so an item with no families at all (no work, no result)
scores log(1 / ITEM_MS0) and ranks behind every item that yields more than
one result per ITEM_MS0 of lane time.  The library then takes the items with
the best prediction first; the prediction never changes what an item yields,
only when it runs.
"""
from __future__ import annotations

import numpy as np

ITEM_MS0 = 10.0  # lane ms added to every item's cost (a probe of 1024 ranks takes 10-200)


def item_target(results: np.ndarray, lane_ms: np.ndarray) -> np.ndarray:
    return np.log((np.asarray(results, dtype=np.float64) + 1.0) /
                  (np.asarray(lane_ms, dtype=np.float64) + ITEM_MS0)).astype(np.float32)


def probe_samples(engine, k: int, first: int, count: int, ranks: int = 1024, idle_ratio: float = 1.0,
                  seed: int = 1):
    """Features and targets of items first .. first+count-1 of the engine's
    graph.  Items with no work (no family) are many on power-law graphs and all
    alike; at most idle_ratio times as many of them as working items are kept
    (a seeded sample)."""
    res, ms = engine.probe_items(first, count, k, ranks)
    X = engine.item_features(first, count, k)
    y = item_target(res, ms)
    work = np.flatnonzero(ms > 0)
    idle = np.flatnonzero(ms <= 0)
    keep_idle = min(len(idle), max(1, int(idle_ratio * len(work))))
    rng = np.random.default_rng(seed)
    pick = np.concatenate([work, rng.choice(idle, keep_idle, replace=False) if len(idle) else idle])
    pick.sort()
    return X[pick], y[pick], {"items": count, "working": int(len(work)), "results": int(res.sum()),
                              "lane_ms": float(ms.sum())}


def train_item_model(engine, samples, epochs: int = 400, lr: float = 3e-3, seed: int = 1):
    """samples: list of (X, y) from probe_samples; returns (model, loss)."""
    X = np.concatenate([s[0] for s in samples]).astype(np.float32)
    y = np.concatenate([s[1] for s in samples]).astype(np.float32)
    return engine.mlp_train(X, y, epochs=epochs, lr=lr, seed=seed)
