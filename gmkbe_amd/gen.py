"""Seeded synthetic bipartite graphs for the measured workloads.

Every generator returns raw edge arrays (u in [0, nU), v in [nU, nU + nV)) that go
through the converter (``convert.convert_edges``), so the relabel path is always
exercised, exactly as a user's edge list would be.  numpy's PCG64 stream is
fixed for a given seed, so the same call yields the same graph on every machine
with this numpy.

Workloads (sizes are the ones the benchmarks and tests use):
  * er(200, 300, 0.05)                      small Erdos-Renyi graph, k = 1
  * chung_lu(50_000, 30_000, 500_000, 2.1, cap=5_000)
  * er_batch(2_000 graphs, sides log-uniform in [20, 2000], p log-uniform in
    [0.01, 0.2], k uniform in {1, 2})
  * chung_lu(2_000_000, 1_000_000, 20_000_000, 2.0)
  * planted(500_000, 500_000, 5_000_000, 2.2, 200 near-bicliques of sides
    uniform in [8, 40], at most 3 missing edges per vertex)
  * hub(n_u, n_v, ...)                      a hub adjacent to every left vertex:
    seed partitions with 4,500 to 40,000 vertices on a side (wide bitsets)
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .convert import convert_edges
from .graphio import Graph

SEED_BASE = 0x6D6B4250_00000000


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def er_edges(n_u: int, n_v: int, p: float, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Each of the n_u * n_v pairs is kept independently with probability p."""
    rng = _rng(seed)
    us, vs = [], []
    rows = max(1, (1 << 22) // max(n_v, 1))
    for r0 in range(0, n_u, rows):
        r1 = min(n_u, r0 + rows)
        m = rng.random((r1 - r0, n_v)) < p
        a, b = np.nonzero(m)
        us.append(a + r0)
        vs.append(b)
    u = np.concatenate(us).astype(np.uint32)
    v = (np.concatenate(vs) + n_u).astype(np.uint32)
    return u, v


def pareto_weights(n: int, alpha: float, total: float, cap: float | None, rng) -> np.ndarray:
    """Expected degrees with tail P(W > x) ~ x^-(alpha - 1), scaled to sum to
    `total`, then capped."""
    x = (1.0 - rng.random(n)) ** (-1.0 / (alpha - 1.0))
    x *= total / x.sum()
    if cap is not None:
        x = np.minimum(x, cap)
    return x


def _sample_cdf(cdf: np.ndarray, x: np.ndarray) -> np.ndarray:
    """searchsorted(cdf, x, side='right'), split over threads (numpy releases
    the GIL inside the search)."""
    n = x.size
    parts = min(16, max(1, (os.cpu_count() or 1)), max(1, n >> 20))
    if parts == 1:
        return np.searchsorted(cdf, x, side="right")
    out = np.empty(n, dtype=np.int64)
    bounds = np.linspace(0, n, parts + 1).astype(np.int64)

    def work(i):
        lo, hi = bounds[i], bounds[i + 1]
        out[lo:hi] = np.searchsorted(cdf, x[lo:hi], side="right")

    with ThreadPoolExecutor(parts) as ex:
        list(ex.map(work, range(parts)))
    return out


def chung_lu_edges(n_u: int, n_v: int, m: int, alpha: float, seed: int, cap: float | None = None,
                   alpha_v: float | None = None) -> tuple[np.ndarray, np.ndarray]:
    """Chung-Lu style bipartite graph: endpoints drawn independently with
    probability proportional to Pareto expected degrees on each side, repeated
    draws collapsed, until m distinct edges exist (first m kept in draw order).

    Draws come in rounds of 1.15x the edges still missing (+1024); a round's
    keys are deduplicated within the round (first occurrence) and against a
    sorted set of the keys kept so far."""
    rng = _rng(seed)
    wu = pareto_weights(n_u, alpha, m, cap, rng)
    wv = pareto_weights(n_v, alpha if alpha_v is None else alpha_v, m, cap, rng)
    cu = np.cumsum(wu); cu /= cu[-1]
    cv = np.cumsum(wv); cv /= cv[-1]
    seen = np.empty(0, dtype=np.int64)  # sorted keys kept so far
    kept = []
    total = 0
    need = m
    for _ in range(64):
        batch = int(need * 1.15) + 1024
        a = np.minimum(_sample_cdf(cu, rng.random(batch)), n_u - 1)
        b = np.minimum(_sample_cdf(cv, rng.random(batch)), n_v - 1)
        k = a * n_v + b
        _, first = np.unique(k, return_index=True)
        first.sort()
        k = k[first]
        if seen.size:
            pos = np.minimum(np.searchsorted(seen, k), seen.size - 1)
            k = k[seen[pos] != k]
        kept.append(k)
        total += k.size
        if total >= m:
            break
        seen = np.sort(np.concatenate([seen, k]), kind="stable")
        need = m - total
    keys = np.concatenate(kept)[:m]
    u = (keys // n_v).astype(np.uint32)
    v = (keys % n_v + n_u).astype(np.uint32)
    return u, v


def planted_edges(n_u: int, n_v: int, m: int, alpha: float, n_cores: int, side_lo: int, side_hi: int,
                  max_miss: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Chung-Lu background plus planted near-bicliques on disjoint vertex sets;
    each planted biclique then loses random edges with at most `max_miss` removed
    per vertex."""
    u, v = chung_lu_edges(n_u, n_v, m, alpha, seed)
    rng = _rng(seed ^ 0x5A5A5A5A)
    perm_u = rng.permutation(n_u)
    perm_v = rng.permutation(n_v)
    pu = pv = 0
    extra_u, extra_v = [], []
    removed = set()
    for _ in range(n_cores):
        a = int(rng.integers(side_lo, side_hi + 1))
        b = int(rng.integers(side_lo, side_hi + 1))
        A = perm_u[pu:pu + a]; pu += a
        B = perm_v[pv:pv + b]; pv += b
        cnt_b = {int(y): 0 for y in B}
        for x in A:
            t = int(rng.integers(0, max_miss + 1))
            cand = [int(y) for y in rng.permutation(B) if cnt_b[int(y)] < max_miss][:t]
            for y in cand:
                cnt_b[y] += 1
                removed.add((int(x), y))
        for x in A:
            for y in B:
                if (int(x), int(y)) not in removed:
                    extra_u.append(int(x)); extra_v.append(int(y))
    # background edges inside a planted block that were removed stay removed
    uu = np.concatenate([u.astype(np.int64) , np.asarray(extra_u, dtype=np.int64)])
    vv = np.concatenate([v.astype(np.int64) - n_u, np.asarray(extra_v, dtype=np.int64)])
    if removed:
        rk = np.asarray([x * n_v + y for x, y in removed], dtype=np.int64)
        keep = ~np.isin(uu * n_v + vv, rk)
        uu, vv = uu[keep], vv[keep]
    return uu.astype(np.uint32), (vv + n_u).astype(np.uint32)


def hub_edges(n_u: int, n_v: int, blocks: int, block_side: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """A graph whose seed partitions are wide: right vertex 0 (the hub) is
    adjacent to every left vertex, so the 2-hop set of any left seed is the
    whole left side; every left vertex also gets one random right neighbour,
    and `blocks` planted block_side x block_side bicliques (on random vertex
    sets) each lose about 10% of their edges."""
    rng = _rng(seed)
    us = [np.arange(n_u, dtype=np.int64), np.arange(n_u, dtype=np.int64)]
    vs = [np.zeros(n_u, dtype=np.int64), rng.integers(1, n_v, n_u)]
    for _ in range(blocks):
        A = rng.choice(n_u, block_side, replace=False)
        B = rng.choice(np.arange(1, n_v), block_side, replace=False)
        a, b = np.meshgrid(A, B, indexing="ij")
        keep = rng.random(a.size) > 0.1
        us.append(a.ravel()[keep]); vs.append(b.ravel()[keep])
    e = np.unique(np.stack([np.concatenate(us), np.concatenate(vs)], 1), axis=0)
    return e[:, 0].astype(np.uint32), (e[:, 1] + n_u).astype(np.uint32)


def hub(n_u: int, n_v: int, blocks: int, block_side: int, seed: int) -> Graph:
    return convert_edges(*hub_edges(n_u, n_v, blocks, block_side, seed))


def er(n_u: int, n_v: int, p: float, seed: int) -> Graph:
    return convert_edges(*er_edges(n_u, n_v, p, seed))


def chung_lu(n_u: int, n_v: int, m: int, alpha: float, seed: int, cap: float | None = None) -> Graph:
    return convert_edges(*chung_lu_edges(n_u, n_v, m, alpha, seed, cap))


def planted(n_u: int, n_v: int, m: int, alpha: float, seed: int, n_cores: int = 200, side_lo: int = 8,
            side_hi: int = 40, max_miss: int = 3) -> Graph:
    return convert_edges(*planted_edges(n_u, n_v, m, alpha, n_cores, side_lo, side_hi, max_miss, seed))


def er_batch_specs(count: int, seed: int) -> list[tuple[int, int, float, int, int]]:
    """(nU, nV, p, k, seed) for each graph of the many-small-graphs workload."""
    rng = _rng(seed)
    out = []
    for i in range(count):
        nu = int(round(np.exp(rng.uniform(np.log(20), np.log(2000)))))
        nv = int(round(np.exp(rng.uniform(np.log(20), np.log(2000)))))
        p = float(np.exp(rng.uniform(np.log(0.01), np.log(0.2))))
        k = int(rng.integers(1, 3))
        out.append((nu, nv, p, k, seed + 1 + i))
    return out


# Seeds of the named workloads.
SEED_C1 = SEED_BASE + 1
SEED_C2 = SEED_BASE + 2
SEED_C3 = SEED_BASE + 3
SEED_C4 = SEED_BASE + 4
SEED_C5 = SEED_BASE + 5


def workload(name: str) -> Graph:
    if name == "er200x300":
        return er(200, 300, 0.05, SEED_C1)
    if name == "cl50kx30k":
        return chung_lu(50_000, 30_000, 500_000, 2.1, SEED_C2, cap=5_000)
    if name == "cl2Mx1M":
        return chung_lu(2_000_000, 1_000_000, 20_000_000, 2.0, SEED_C4)
    if name == "planted500k":
        return planted(500_000, 500_000, 5_000_000, 2.2, SEED_C5)
    raise KeyError(name)
