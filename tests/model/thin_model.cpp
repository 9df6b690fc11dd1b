// CPU model of the thin-result kernels: runs the same family construction
// (thin_build.h) and maximality test (thin.h) sequentially, over all tuples of
// one side.  Test infrastructure only: it lets the CPU tests check that
// thin-L + thin-R + seed partitions cover exactly the reference's result set.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../gmkbe_amd/csrc/thin.h"
#include "../../gmkbe_amd/csrc/thin_build.h"

using namespace gmk::thin;

namespace {
// Arena with stable addresses (families keep pointers into it).
struct HostArena {
  std::vector<std::vector<u32>> blocks;
  u64 alloc(u64 n) { blocks.emplace_back(n ? n : 1); return blocks.size() - 1; }
  u32* ptr(u64 id) { return blocks[id].data(); }
};
struct HostFams {
  std::vector<Family> f;
  Family* next() { f.emplace_back(); return &f.back(); }
};

// Neighbour hashes of the vertices of degree >= nh_min, in the layout the
// device builds (thin.h): sequential inserts with the same slot function.
struct HostHashes {
  std::vector<u64> off;
  std::vector<u32> slots;
  void build(unsigned n, const unsigned* goff, const unsigned* adj, u32 nh_min) {
    off.assign(n, 0);
    u64 total = 0;
    for (u32 v = 0; v < n; ++v) {
      const u32 d = goff[v + 1] - goff[v];
      if (d >= nh_min) { off[v] = total; total += 1ull << nh_log2(d); }
    }
    slots.assign(total ? total : 1, NH_EMPTY);
    for (u32 v = 0; v < n; ++v) {
      const u32 d = goff[v + 1] - goff[v];
      if (d < nh_min) continue;
      const u32 lg = nh_log2(d), mask = (1u << lg) - 1u;
      u32* t = slots.data() + off[v];
      for (u32 e = goff[v]; e < goff[v + 1]; ++e) {
        u32 s = nh_slot(adj[e], lg);
        while (t[s] != NH_EMPTY) s = (s + 1) & mask;
        t[s] = adj[e];
      }
    }
  }
};
u32 g_hash_min = 0;  // 0: the model uses sorted-list searches only
bool g_guards = true;  // guard tables for the checks, as the device builds them
unsigned long long g_modes[4] = {0, 0, 0, 0};  // families per guard mode, and dropped as dominated
}  // namespace

// Families of the thin_model calls since the last read, by guard mode
// (GUARD_NONE, GUARD_TABLE, GUARD_ALLMAX) and dropped as dominated (index 3).
extern "C" void thin_model_guard_modes(unsigned long long* out) {
  for (int i = 0; i < 4; ++i) { out[i] = g_modes[i]; g_modes[i] = 0; }
}

// Later thin_model calls build the families' guard tables (1, the default) or
// check every rank with the plain candidate lists (0).
extern "C" void thin_model_set_guards(int on) { g_guards = on != 0; }

// Later thin_model calls use neighbour hashes for vertices of degree >= v (0: none).
extern "C" void thin_model_set_hash_min(unsigned v) { g_hash_min = v; }

// Membership through the neighbour hashes against the sorted lists, for every
// hashed vertex v and every vertex x: the number of disagreements.
extern "C" long long thin_model_hash_check(unsigned n, const unsigned* off, const unsigned* adj, unsigned nh_min) {
  HostHashes h;
  h.build(n, off, adj, nh_min);
  Graph g;
  g.off = off; g.adj = adj; g.nh_off = h.off.data(); g.nh = h.slots.data(); g.nh_min = nh_min;
  long long bad = 0;
  Cost cost;
  for (u32 v = 0; v < n; ++v) {
    const u32 d = off[v + 1] - off[v];
    const u32* t = nh_table(g, v, d, cost);
    if (!t) continue;
    for (u32 x = 0; x < n; ++x)
      bad += nh_find(t, d, x, cost) != contains(adj + off[v], d, x, cost);
  }
  return bad;
}

extern "C" long long thin_model(unsigned n, unsigned n_left, const unsigned* off, const unsigned* adj, unsigned k,
                                unsigned theta, int side_a, unsigned min_b, unsigned** out_words,
                                unsigned long long* out_n) {
  Graph g;
  g.off = off; g.adj = adj;
  if (side_a == 0) { g.a_lo = 0; g.a_hi = n_left; g.b_lo = n_left; g.b_hi = n; }
  else { g.a_lo = n_left; g.a_hi = n; g.b_lo = 0; g.b_hi = n_left; }
  HostHashes hh;
  if (g_hash_min) {
    hh.build(n, off, adj, g_hash_min);
    g.nh_off = hh.off.data(); g.nh = hh.slots.data(); g.nh_min = g_hash_min;
  }
  std::vector<u32> words;
  long long count = 0;
  const u32 jmin = theta > 1 ? theta : 1;
  for (u32 s = g.a_lo; s < g.a_hi; ++s) {
    for (u32 j = jmin; j <= 2 * k && j <= (u32)MAXJ; ++j) {
      // all tuples (s, w_1 < ... < w_{j-1}) with w_i > s
      u32 L[MAXJ];
      L[0] = s;
      std::vector<u32> idx(j, 0);
      for (u32 i = 1; i < j; ++i) idx[i] = s + i;
      if (j > 1 && idx[j - 1] >= g.a_hi) continue;
      for (;;) {
        for (u32 i = 1; i < j; ++i) L[i] = idx[i];
        HostArena ar;
        HostFams fo;
        Cost cost;
        build_families(g, L, j, k, min_b, 0, ar, fo, cost);
        if (g_guards)
          for (Family& f : fo.f) {
            const bool live = f.total > 0;
            guard_build(g, f, k, ar, cost, 0);  // every family: the tests exercise the tables
            if (live && f.total == 0) ++g_modes[3];
            else if (live) ++g_modes[f.gmode];
          }
        for (const Family& f : fo.f) {
          for (u64 r = 0; r < f.total; ++r) {
            u64 rr = r;
            u32 ch[MAXJ * 3], ct[MAXJ * 3], nch = 0;
            for (u32 t = 0; t < f.ntypes; ++t) {
              const u64 radix = binom(f.psize[t], f.cnt[t]);
              const u64 rt = rr % radix;
              rr /= radix;
              u32 id[3];
              unrank(rt, f.cnt[t], f.psize[t], id);
              for (u32 q = 0; q < f.cnt[t]; ++q) {
                ch[nch] = pool_member(f, t, id[q], g.b_lo, cost);
                ct[nch] = popc(f.mask[t]);
                ++nch;
              }
            }
            if (side_a_maximal<MAXJ * 3>(g, fam_ref(f), ch, ct, nch, k, cost) != 1)
              continue;
            std::vector<u32> R(f.freep, f.freep + f.nfree);
            R.insert(R.end(), ch, ch + nch);
            std::sort(R.begin(), R.end());
            std::vector<u32> T(f.L, f.L + f.j);
            const std::vector<u32>& left = side_a == 0 ? T : R;
            const std::vector<u32>& right = side_a == 0 ? R : T;
            words.push_back((u32)left.size());
            words.insert(words.end(), left.begin(), left.end());
            words.push_back((u32)right.size());
            words.insert(words.end(), right.begin(), right.end());
            ++count;
          }
        }
        // next tuple
        if (j == 1) break;
        int p = (int)j - 1;
        while (p >= 1 && idx[p] == g.a_hi - (j - p)) --p;
        if (p < 1) break;
        ++idx[p];
        for (u32 q = p + 1; q < j; ++q) idx[q] = idx[q - 1] + 1;
      }
    }
  }
  *out_n = words.size();
  *out_words = (u32*)malloc(sizeof(u32) * (words.size() + 1));
  if (!words.empty()) memcpy(*out_words, words.data(), sizeof(u32) * words.size());
  return count;
}

// Number of one-vertex tuples of side_a with a dominating vertex: some w outside
// the tuple misses at most k of the tuple's free list (dominance_bound <= k).
// Lets the tests check that graphs meant to exercise the family skip do.
extern "C" long long thin_dominated(unsigned n, unsigned n_left, const unsigned* off, const unsigned* adj, unsigned k,
                                    int side_a) {
  Graph g;
  g.off = off; g.adj = adj;
  if (side_a == 0) { g.a_lo = 0; g.a_hi = n_left; g.b_lo = n_left; g.b_hi = n; }
  else { g.a_lo = n_left; g.a_hi = n; g.b_lo = 0; g.b_hi = n_left; }
  long long c = 0;
  for (u32 s = g.a_lo; s < g.a_hi; ++s) {
    Cost cost;
    const u32 L[1] = {s};
    if (dominance_bound(g, L, 1, k, adj + off[s], off[s + 1] - off[s], 0, 1, cost) <= k) ++c;
  }
  return c;
}

// Prefix sums of family totals by the phases of the device scan
// (thin_scan_totals: thin::scan_phase_*), run thread by thread, each barrier of
// the block becoming the end of a loop over the TSCAN threads.  prefix gets
// n + 1 entries.
extern "C" void thin_scan_totals_model(const unsigned long long* totals, unsigned n, unsigned long long* prefix) {
  std::vector<u64> part(TSCAN), next(TSCAN);
  auto total = [totals](u32 i) { return (u64)totals[i]; };
  for (u32 t = 0; t < TSCAN; ++t) part[t] = scan_phase_sum(n, t, total);
  for (u32 d = 1; d < TSCAN; d <<= 1) {
    for (u32 t = 0; t < TSCAN; ++t) next[t] = scan_phase_step(part.data(), t, d);
    part.swap(next);
  }
  for (u32 t = 0; t < TSCAN; ++t) scan_phase_write(n, t, part.data(), total, (u64*)prefix);
}

// The device's combination arithmetic on the host: C(n, c) saturating at
// 2^62, and the colex unrank of rank r of C(n, c) into out[0 .. c-1].
extern "C" unsigned long long thin_model_binom(unsigned long long n, unsigned c) { return binom(n, c); }
extern "C" void thin_model_unrank(unsigned long long r, unsigned c, unsigned long long n, unsigned* out) {
  u32 idx[3] = {0, 0, 0};
  unrank(r, c, n, idx);
  for (u32 i = 0; i < c && i < 3; ++i) out[i] = idx[i];
}
