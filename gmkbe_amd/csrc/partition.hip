// Partition construction on the host: the global partition of a small graph and
// the seed partitions, plus the seeded-mode driver.
#include "gmk_host.h"
#include "seeds.h"
#include "thin.h"

#include <algorithm>
#include <string>

#include <cstdlib>
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

SeedJob gmk::make_seed_job(const HostGraph& g, u32 k, u32 lb, u32 graph) {
  SeedJob j;
  j.g = g;
  j.k = k;
  j.lb = lb;
  j.graph = graph;
  degeneracy_order(g, j.order);
  u32 maxd[2] = {0, 0};
  for (u32 v = 0; v < g.n; ++v) maxd[v >= g.n_left] = std::max(maxd[v >= g.n_left], g.deg(v));
  const int us = maxd[0] < maxd[1] ? 0 : 1;
  const u32 lo = us ? g.n_left : 0, hi = us ? g.n : g.n_left;
  for (u32 v = lo; v < hi; ++v)
    if (g.deg(v) > 0) j.seeds.push_back(v);  // a vertex without neighbours seeds no result with both sides >= 1
  std::sort(j.seeds.begin(), j.seeds.end(), [&](u32 a, u32 b) { return j.order[a] < j.order[b]; });
  return j;
}

// Seeded mode: thin results first (tuple side left, then right), then the fat
// results from seed partitions, seeds taken in degeneracy order.
int gmk_run_seeded(gmk_ctx* c, const gmk_params* prm, gmk_stats* st, RunState& rs) {
  const u32 k = prm->k, theta = prm->theta;
  int r = GMK_OK;
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
  // the fat results live in the (lb - k)-core (BPPivot's core reduction, applied
  // when lb > k: biplex.cpp:441-450); it is built on the device
  CoreGraph core, bfly;
  if ((r = core_reduce(c, lb - k, rs, core))) return r;
  // then BPPivot's butterfly reduction, applied when lb > 2k (biplex.cpp:469-478),
  // as always here; GMK_BUTTERFLY=0 leaves it out
  const CoreGraph* red = &core;
  const char* bf = getenv("GMK_BUTTERFLY");
  if (!bf || atoi(bf) != 0) {
    if ((r = butterfly_reduce(c, core, lb, k, rs, bfly))) return r;
    red = &bfly;
  }
  HostGraph g{c->n, c->n_left, red->off.data(), red->adj.data()};
  std::vector<SeedJob> jobs;
  jobs.push_back(make_seed_job(g, k, lb, 0));
  jobs[0].d_off = red->d_off;
  jobs[0].d_adj = red->d_adj;
  return run_seed_jobs(c, jobs, prm, st, rs, nullptr);
}
