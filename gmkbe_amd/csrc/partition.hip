// Partition construction on the host: the global partition of a small graph.
// Seed partitions are built on the device (seed_kernels.hip).
#include "gmk_host.h"

#include <cstring>

namespace gmk {

u64* push_root(PartitionSet& ps, u32 pid) {
  const size_t stride = 1 + 6 * (size_t)ps.W;
  const size_t at = ps.roots.size();
  ps.roots.resize(at + stride, 0);
  ps.roots[at] = pid;
  return ps.roots.data() + at + 1;
}

void build_global_partition(u32 n, u32 n_left, const u32* off, const u32* adj, u32 theta, int W, PartitionSet& ps) {
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
  const u32 pid = (u32)ps.parts.size();
  ps.parts.push_back(P);
  u64* st = push_root(ps, pid);
  // C_L = all left, C_R = all right
  for (u32 u = 0; u < n_left; ++u) st[1 * W + u / 64] |= 1ull << (u % 64);
  for (u32 v = 0; v < nr; ++v) st[4 * W + v / 64] |= 1ull << (v % 64);
}

}  // namespace gmk

int gmk_run_seeded(gmk_ctx* c, const gmk_params* prm, gmk_stats* st, gmk::RunState& rs) {
  (void)c; (void)prm; (void)st; (void)rs;
  return GMK_EUNSUPPORTED;
}
