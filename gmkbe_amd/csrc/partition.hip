// Partition construction on the host: the global partition of a small graph and
// the seed partitions, plus the seeded-mode driver.
#include "gmk_host.h"
#include "seeds.h"
#include "thin.h"

#include <algorithm>
#include <string>

#include <cstring>

namespace gmk {

u64* push_root(PartitionSet& ps, u32 pid) {
  const size_t stride = 1 + 6 * (size_t)ps.W;
  const size_t at = ps.roots.size();
  ps.roots.resize(at + stride, 0);
  ps.roots[at] = pid;
  return ps.roots.data() + at + 1;
}

void build_global_partition(u32 n, u32 n_left, const u32* off, const u32* adj, u32 k, u32 theta, int W, PartitionSet& ps,
                            u32 graph) {
  const u32 nr = n - n_left;
  const u32 wl = (n_left + 63) / 64, wr = (nr + 63) / 64;
  ps.W = W;
  PartDesc P{};
  P.nl = n_left; P.nr = nr; P.wl = wl; P.wr = wr;
  P.rows_l = ps.rows.size();
  ps.rows.resize(ps.rows.size() + (size_t)n_left * wr, 0);
  for (u32 u = 0; u < n_left; ++u)
    for (u32 e = off[u]; e < off[u + 1]; ++e) {
      const u32 v = adj[e] - n_left;
      ps.rows[P.rows_l + (size_t)u * wr + v / 64] |= 1ull << (v % 64);
    }
  P.rows_r = ps.rows.size();
  ps.rows.resize(ps.rows.size() + (size_t)nr * wl, 0);
  for (u32 v = 0; v < nr; ++v)
    for (u32 e = off[n_left + v]; e < off[n_left + v + 1]; ++e) {
      const u32 u = adj[e];
      ps.rows[P.rows_r + (size_t)v * wl + u / 64] |= 1ull << (u % 64);
    }
  P.map_l = ps.maps.size();
  for (u32 u = 0; u < n_left; ++u) ps.maps.push_back(u);
  P.map_r = ps.maps.size();
  for (u32 v = 0; v < nr; ++v) ps.maps.push_back(n_left + v);
  P.lb = theta;
  P.seed = NONE;
  P.k = k;
  P.graph = graph;
  const u32 pid = (u32)ps.parts.size();
  ps.parts.push_back(P);
  u64* st = push_root(ps, pid);
  // C_L = all left, C_R = all right
  for (u32 u = 0; u < n_left; ++u) st[1 * W + u / 64] |= 1ull << (u % 64);
  for (u32 v = 0; v < nr; ++v) st[4 * W + v / 64] |= 1ull << (v % 64);
}

}  // namespace gmk

using namespace gmk;

// Seeded mode: thin results first (tuple side left, then right), then the fat
// results from seed partitions, seeds taken in degeneracy order.
int gmk_run_seeded(gmk_ctx* c, const gmk_params* prm, gmk_stats* st, RunState& rs) {
  const u32 k = prm->k, theta = prm->theta;
  int r;
  if (theta <= 2 * k) {
    if (2 * k > (u32)thin::MAXJ) {
      c->err = "seeded mode with theta <= 2k supports k <= 3 (use the global partition)";
      return GMK_EUNSUPPORTED;
    }
    if ((r = gmk_run_thin(c, prm, st, rs, 0, theta))) return r;
    if (rs.hit_limit) return GMK_OK;
    if ((r = gmk_run_thin(c, prm, st, rs, 1, std::max(theta, 2 * k + 1)))) return r;
    if (rs.hit_limit) return GMK_OK;
  }
  const u32 lb = std::max(theta, 2 * k + 1);
  HostGraph g{c->n, c->n_left, c->off.data(), c->adj.data()};
  if (!c->order_valid) { degeneracy_order(g, c->order); c->order_valid = true; }
  u32 maxd[2] = {0, 0};
  for (u32 v = 0; v < c->n; ++v) maxd[v >= c->n_left] = std::max(maxd[v >= c->n_left], g.deg(v));
  const int us = maxd[0] < maxd[1] ? 0 : 1;
  const u32 lo = us ? c->n_left : 0, hi = us ? c->n : c->n_left;
  std::vector<u32> seeds;
  for (u32 v = lo; v < hi; ++v) seeds.push_back(v);
  std::sort(seeds.begin(), seeds.end(), [&](u32 a, u32 b) { return c->order[a] < c->order[b]; });
  std::vector<int32_t> cnt(c->n, 0);
  std::vector<uint8_t> mark(c->n, 0);
  std::vector<u32> loc(c->n, NONE);
  // partitions are grouped by bitset width; a group is run when it is full
  const size_t group_cap = prm->mode == GMK_MODE_PLAIN ? 1 : 4096;
  PartitionSet groups[7];
  size_t gcount[7] = {0};
  auto flush = [&](int gi) -> int {
    if (gcount[gi] == 0) return GMK_OK;
    int rr = gmk_run_bb(c, groups[gi], prm, st, rs);
    st->partitions += gcount[gi];
    groups[gi] = PartitionSet();
    gcount[gi] = 0;
    return rr;
  };
  SeedPartition sp;
  for (u32 s : seeds) {
    if (!build_seed_partition(g, c->order, s, k, lb, sp, cnt, mark)) continue;
    const u32 nl = (u32)sp.left.size(), nr = (u32)sp.right.size();
    const u32 wl = (nl + 63) / 64, wr = (nr + 63) / 64;
    const int W = gmk_pick_w(std::max(wl, wr));
    if (!W) {
      c->err = "seed partition of vertex " + std::to_string(s) + " has " + std::to_string(std::max(nl, nr)) +
               " vertices on one side (limit 4096)";
      return GMK_EUNSUPPORTED;
    }
    int gi = 0;
    while ((1 << gi) < W) ++gi;
    PartitionSet& ps = groups[gi];
    ps.W = W;
    PartDesc P{};
    P.nl = nl; P.nr = nr; P.wl = wl; P.wr = wr;
    P.lb = lb;
    P.seed = s;
    P.k = k;
    P.graph = 0;
    P.map_l = ps.maps.size();
    ps.maps.insert(ps.maps.end(), sp.left.begin(), sp.left.end());
    P.map_r = ps.maps.size();
    ps.maps.insert(ps.maps.end(), sp.right.begin(), sp.right.end());
    for (u32 i = 0; i < nl; ++i) loc[sp.left[i]] = i;
    for (u32 i = 0; i < nr; ++i) loc[sp.right[i]] = i;
    P.rows_l = ps.rows.size();
    ps.rows.resize(ps.rows.size() + (size_t)nl * wr, 0);
    for (u32 i = 0; i < nl; ++i) {
      const u32 u = sp.left[i];
      for (u32 e = g.off[u]; e < g.off[u + 1]; ++e) {
        const u32 q = loc[g.adj[e]];
        if (q != NONE) ps.rows[P.rows_l + (size_t)i * wr + q / 64] |= 1ull << (q % 64);
      }
    }
    P.rows_r = ps.rows.size();
    ps.rows.resize(ps.rows.size() + (size_t)nr * wl, 0);
    for (u32 i = 0; i < nr; ++i) {
      const u32 v = sp.right[i];
      for (u32 e = g.off[v]; e < g.off[v + 1]; ++e) {
        const u32 q = loc[g.adj[e]];
        if (q != NONE) ps.rows[P.rows_r + (size_t)i * wl + q / 64] |= 1ull << (q % 64);
      }
    }
    for (u32 v : sp.left) loc[v] = NONE;
    for (u32 v : sp.right) loc[v] = NONE;
    const u32 pid = (u32)ps.parts.size();
    ps.parts.push_back(P);
    u64* root = push_root(ps, pid);
    // root: S = {seed}, C / X from the flags (0 = C, 1 = X, 2 = S)
    for (int side = 0; side < 2; ++side) {
      const std::vector<uint8_t>& fl = side ? sp.rflag : sp.lflag;
      for (u32 i = 0; i < fl.size(); ++i) {
        const int which = fl[i] == 2 ? 0 : (fl[i] == 0 ? 1 : 2);
        root[(side * 3 + which) * W + i / 64] |= 1ull << (i % 64);
      }
    }
    if (++gcount[gi] >= group_cap) {
      if ((r = flush(gi))) return r;
      if (rs.hit_limit) return GMK_OK;
    }
  }
  for (int gi = 0; gi < 7; ++gi) {
    if ((r = flush(gi))) return r;
    if (rs.hit_limit) return GMK_OK;
  }
  return GMK_OK;
}
