// Host construction of seed partitions (see seeds.h).
#include "seeds.h"

#include <algorithm>

namespace gmk {

void degeneracy_order(const HostGraph& g, std::vector<uint32_t>& order) {
  const uint32_t n = g.n;
  std::vector<uint32_t> deg(n), pos(n), vert(n);
  uint32_t maxd = 0;
  for (uint32_t v = 0; v < n; ++v) { deg[v] = g.deg(v); maxd = std::max(maxd, deg[v]); }
  // bin sort by degree (Batagelj-Zaversnik), ties by vertex ID
  std::vector<uint32_t> bin(maxd + 2, 0);
  for (uint32_t v = 0; v < n; ++v) ++bin[deg[v] + 1];
  for (uint32_t d = 1; d <= maxd + 1; ++d) bin[d] += bin[d - 1];
  std::vector<uint32_t> start(bin.begin(), bin.end() - 1);
  for (uint32_t v = 0; v < n; ++v) { pos[v] = start[deg[v]]++; vert[pos[v]] = v; }
  // bin[d] = first position of degree d
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t v = vert[i];
    for (uint32_t e = g.off[v]; e < g.off[v + 1]; ++e) {
      const uint32_t w = g.adj[e];
      if (deg[w] > deg[v]) {
        const uint32_t dw = deg[w], pw = pos[w], ps = bin[dw], u = vert[ps];
        if (u != w) { pos[u] = pw; vert[pw] = u; pos[w] = ps; vert[ps] = w; }
        ++bin[dw];
        --deg[w];
      }
    }
  }
  order.resize(n);
  for (uint32_t i = 0; i < n; ++i) order[vert[i]] = i;
}

namespace {
enum : uint8_t { M_NONE = 0, M_CS = 1, M_XS = 2, M_CO_N = 3, M_SEED = 4, M_CO_3 = 5, M_DEAD = 6, M_SEEN3 = 7 };
}

bool build_seed_partition(const HostGraph& g, const std::vector<uint32_t>& order, uint32_t s, uint32_t k,
                          uint32_t lb, SeedPartition& out, std::vector<int32_t>& cnt, std::vector<uint8_t>& mark) {
  std::vector<uint32_t> CS, XS, CO, touched;
  auto touch = [&](uint32_t v, uint8_t m) { if (mark[v] == M_NONE) touched.push_back(v); mark[v] = m; };
  touch(s, M_SEED);
  for (uint32_t e = g.off[s]; e < g.off[s + 1]; ++e) { const uint32_t v = g.adj[e]; touch(v, M_CO_N); CO.push_back(v); }
  for (uint32_t v : CO)
    for (uint32_t e = g.off[v]; e < g.off[v + 1]; ++e) {
      const uint32_t w = g.adj[e];
      if (mark[w] != M_NONE) continue;
      if (order[w] > order[s]) { touch(w, M_CS); CS.push_back(w); }
      else { touch(w, M_XS); XS.push_back(w); }
    }
  auto same_side_member = [&](uint8_t m) { return m == M_CS || m == M_XS || m == M_SEED; };
  auto is_co = [&](uint8_t m) { return m == M_CO_N || m == M_CO_3; };
  // cnt[x] (x on the seed's side) = neighbours in C_os; cnt[y] (y in C_os) = neighbours in C_us
  for (uint32_t y : CO)
    for (uint32_t e = g.off[y]; e < g.off[y + 1]; ++e) {
      const uint32_t x = g.adj[e];
      if (same_side_member(mark[x])) ++cnt[x];
      if (mark[x] == M_CS) ++cnt[y];
    }
  const int64_t need_os = (int64_t)lb - (int64_t)k;       // C_os vertices: neighbours in L
  const int64_t need_us = (int64_t)lb - 2 * (int64_t)k;   // C_us / X_us vertices: neighbours in C_os
  bool alive = true;
  auto prune = [&]() {
    for (;;) {
      bool changed = false;
      for (uint32_t y : CO) {
        if (!is_co(mark[y])) continue;
        if ((int64_t)cnt[y] + (mark[y] == M_CO_N ? 1 : 0) < need_os) {
          mark[y] = M_DEAD;
          changed = true;
          for (uint32_t e = g.off[y]; e < g.off[y + 1]; ++e) {
            const uint32_t x = g.adj[e];
            if (same_side_member(mark[x])) --cnt[x];
          }
          if ((int64_t)cnt[s] < need_os) return false;
        }
      }
      if (changed) continue;
      for (uint32_t w : CS) {
        if (mark[w] != M_CS || (int64_t)cnt[w] >= need_us) continue;
        mark[w] = M_DEAD;
        changed = true;
        for (uint32_t e = g.off[w]; e < g.off[w + 1]; ++e) {
          const uint32_t y = g.adj[e];
          if (is_co(mark[y])) --cnt[y];
        }
      }
      for (uint32_t x : XS)
        if (mark[x] == M_XS && (int64_t)cnt[x] < need_us) { mark[x] = M_DEAD; changed = true; }
      // the seed needs lb - k neighbours among the surviving C_os (it misses at
      // most k of a side of lb or more); BPPivot's pruneCX tests this only when
      // one of them leaves (biplex.cpp:240-258), so a seed that starts short
      // of it kept a partition that holds no result
      if (!changed) return (int64_t)cnt[s] >= need_os;
    }
  };
  alive = prune();
  if (alive && k >= 1) {
    // 3-hop vertices of the other side: count neighbours in C_us first
    std::vector<uint32_t> cand;
    for (uint32_t w : CS) {
      if (mark[w] != M_CS) continue;
      for (uint32_t e = g.off[w]; e < g.off[w + 1]; ++e) {
        const uint32_t y = g.adj[e];
        if (mark[y] == M_NONE) { touch(y, M_SEEN3); cand.push_back(y); }
        if (mark[y] == M_SEEN3) ++cnt[y];
      }
    }
    for (uint32_t y : cand) {
      if ((int64_t)cnt[y] + (int64_t)k >= (int64_t)lb) {
        mark[y] = M_CO_3;
        CO.push_back(y);
        for (uint32_t e = g.off[y]; e < g.off[y + 1]; ++e) {
          const uint32_t x = g.adj[e];
          if (same_side_member(mark[x])) ++cnt[x];
        }
      } else {
        mark[y] = M_DEAD;
      }
    }
    alive = prune();
  }
  out.seed = s;
  out.us = s < g.n_left ? 0 : 1;
  out.left.clear(); out.right.clear(); out.lflag.clear(); out.rflag.clear();
  if (alive) {
    std::vector<uint32_t> same, other;
    same.push_back(s);
    for (uint32_t w : CS) if (mark[w] == M_CS) same.push_back(w);
    for (uint32_t x : XS) if (mark[x] == M_XS) same.push_back(x);
    for (uint32_t y : CO) if (is_co(mark[y])) other.push_back(y);
    std::sort(same.begin(), same.end());
    std::sort(other.begin(), other.end());
    std::vector<uint8_t> fs(same.size()), fo(other.size(), 0);
    for (size_t i = 0; i < same.size(); ++i) fs[i] = same[i] == s ? 2 : (mark[same[i]] == M_XS ? 1 : 0);
    if (out.us == 0) { out.left = same; out.lflag = fs; out.right = other; out.rflag = fo; }
    else { out.right = same; out.rflag = fs; out.left = other; out.lflag = fo; }
  }
  for (uint32_t v : touched) { mark[v] = M_NONE; cnt[v] = 0; }
  return alive;
}

}  // namespace gmk
