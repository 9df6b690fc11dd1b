// Oracle for maximal k-biplex enumeration.  TEST INFRASTRUCTURE ONLY: nothing in
// the product library links, loads or calls this file; tests, smoke() and the
// CPU-baseline leg of bench.py use it as the checker and as the timed CPU port.
//
// Contents (each function cites the reference code whose behaviour it restates):
//   * or_bppivot      - branch-and-bound enumerator restating BPPivot
//                       (baselines/MaximalBiPlex/src/biplex/biplex.cpp:418-554,
//                       695-1242), global mode and seeded mode.
//   * or_brute        - exhaustive enumeration for tiny graphs (pins the above).
//   * or_validate     - k-biplex / maximality / size / duplicate checker.
//   * or_canon_digest - canonical sorted-set digest and the order-free fingerprint
//                       digest also computed by the GPU path.
// (The .graph.bin converter is gmkbe_amd/convert.py, a product tool; the
// reference converter built under oracle/_ref/ is its checker.)
//
// Conventions.  A graph is the .graph.bin CSR: vertices 0..nL-1 are the left side,
// nL..N-1 the right side, neighbour lists sorted ascending.  A result is recorded as
// u32 words  [|L|, L ascending..., |R|, R ascending...]  with global IDs.
//
// Each function is pinned by tests:
//   or_bppivot  vs the compiled reference binary (oracle/_ref/bppivot) golden
//               vectors and vs or_brute;  or_brute vs closed forms (K_{a,b},
//               matchings, stars).
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

typedef uint32_t u32;
typedef uint64_t u64;

// ---------------------------------------------------------------------------
// Fingerprints.  fp(L,R) = fmix64(sum_{x in L u R} mix(x) + |L|*A + |R|*B): a
// function of the two vertex SETS (order-free), so the GPU can compute it from
// bitsets or unsorted lists.  The set digest is (count, sum fp, xor fp).
// ---------------------------------------------------------------------------
static inline u64 or_mix(u64 x) {  // splitmix64 finaliser
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static inline u64 or_fmix(u64 h) {  // murmur3 fmix64
  h ^= h >> 33; h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33; return h;
}
static const u64 FP_A = 0x2545F4914F6CDD1Dull, FP_B = 0x9E6C63D0676A9A99ull;

static u64 fp_of(const u32* L, u32 nl, const u32* R, u32 nr) {
  u64 s = (u64)nl * FP_A + (u64)nr * FP_B;
  for (u32 i = 0; i < nl; ++i) s += or_mix(L[i]);
  for (u32 i = 0; i < nr; ++i) s += or_mix(R[i]);
  return or_fmix(s);
}

extern "C" {

typedef struct {
  u64 count;
  u64 fp_sum;
  u64 fp_xor;
  u64 canon_lo, canon_hi;  // digest of the sorted canonical record stream
} or_digest;

}  // extern "C"

// Result sink shared by the enumerators.
struct Sink {
  std::vector<u32> words;  // records (may be disabled to count only)
  bool keep = true;
  u64 count = 0, fp_sum = 0, fp_xor = 0;
  u64 limit = ~0ull;
  void emit(const std::vector<u32>& L, const std::vector<u32>& R) {
    std::vector<u32> a(L), b(R);
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    u64 f = fp_of(a.data(), (u32)a.size(), b.data(), (u32)b.size());
    ++count; fp_sum += f; fp_xor ^= f;
    if (keep) {
      words.push_back((u32)a.size());
      words.insert(words.end(), a.begin(), a.end());
      words.push_back((u32)b.size());
      words.insert(words.end(), b.begin(), b.end());
    }
  }
  bool full() const { return count >= limit; }
};

// ---------------------------------------------------------------------------
// Graph view used by the enumerators: side 0 = left, side 1 = right, local IDs.
// ---------------------------------------------------------------------------
struct BiG {
  u32 n[2];
  u32 nL;  // global offset of the right side
  std::vector<std::vector<u32>> nbr[2];  // local IDs of the opposite side, sorted
  std::vector<u64> bits;                 // dense adjacency (left x right) when small
  u32 wR = 0;
  bool dense = false;
  void build(u32 nLeft, u32 N, const u32* off, const u32* adj) {
    nL = nLeft; n[0] = nLeft; n[1] = N - nLeft;
    nbr[0].assign(n[0], {}); nbr[1].assign(n[1], {});
    for (u32 v = 0; v < N; ++v) {
      int s = v < nLeft ? 0 : 1;
      u32 lv = s ? v - nLeft : v;
      for (u32 e = off[v]; e < off[v + 1]; ++e) {
        u32 w = adj[e];
        nbr[s][lv].push_back(s ? w : w - nLeft);
      }
      std::sort(nbr[s][lv].begin(), nbr[s][lv].end());
    }
    wR = (n[1] + 63) / 64;
    if ((u64)n[0] * wR <= (1ull << 24)) {
      dense = true;
      bits.assign((size_t)n[0] * wR, 0);
      for (u32 u = 0; u < n[0]; ++u)
        for (u32 v : nbr[0][u]) bits[(size_t)u * wR + v / 64] |= 1ull << (v % 64);
    }
  }
  bool connect(int side, u32 a, u32 b) const {
    u32 u = side == 0 ? a : b, v = side == 0 ? b : a;  // u left, v right
    if (dense) return (bits[(size_t)u * wR + v / 64] >> (v % 64)) & 1;
    const auto& l = nbr[0][u];
    return std::binary_search(l.begin(), l.end(), v);
  }
};

// ---------------------------------------------------------------------------
// Indexed vertex set with O(1) membership, insertion and removal.  Removals made
// inside a branch are undone from an explicit log (reverse order), which restores
// the same membership; iteration order is not part of the algorithm's contract.
// ---------------------------------------------------------------------------
struct ISet {
  std::vector<u32> items;
  std::vector<int32_t> where;  // -1 when absent
  void init(u32 cap) { items.clear(); where.assign(cap, -1); }
  bool has(u32 v) const { return where[v] >= 0; }
  u32 size() const { return (u32)items.size(); }
  void add(u32 v) { if (where[v] < 0) { where[v] = (int32_t)items.size(); items.push_back(v); } }
  void del(u32 v) {
    int32_t p = where[v];
    if (p < 0) return;
    u32 last = items.back();
    items[p] = last; where[last] = p;
    items.pop_back(); where[v] = -1;
  }
};

// ---------------------------------------------------------------------------
// BPPivot restatement (biplex.cpp:695-1242).  State: S, C (candidates), X
// (excluded) per side; degC[side][v] = |N(v) n C[opp]| for every vertex;
// nonS[side][v] = non-neighbours of v inside S[opp] (at most k+1 kept).
// ---------------------------------------------------------------------------
struct BPP {
  const BiG* G;
  u32 k, lb;
  ISet S[2], C[2], X[2];
  std::vector<u32> degC[2];
  std::vector<std::vector<u32>> nonS[2];
  Sink* sink;
  u64 branches = 0;
  bool stop = false;

  void init_sets() {
    for (int s = 0; s < 2; ++s) {
      S[s].init(G->n[s]); C[s].init(G->n[s]); X[s].init(G->n[s]);
      degC[s].assign(G->n[s], 0);
      nonS[s].assign(G->n[s], {});
    }
  }

  // Removing v (side s) from C lowers degC of its neighbours (biplex.cpp:1084-1096).
  void c_del(int s, u32 v, std::vector<std::pair<int, u32>>& log) {
    if (!C[s].has(v)) return;
    C[s].del(v);
    for (u32 w : G->nbr[s][v]) --degC[s ^ 1][w];
    log.push_back({s, v});
  }
  void c_undo(std::vector<std::pair<int, u32>>& log, size_t mark) {
    while (log.size() > mark) {
      auto [s, v] = log.back(); log.pop_back();
      C[s].add(v);
      for (u32 w : G->nbr[s][v]) ++degC[s ^ 1][w];
    }
  }
  void x_del(int s, u32 v, std::vector<std::pair<int, u32>>& log) {
    if (!X[s].has(v)) return;
    X[s].del(v); log.push_back({s, v});
  }
  void x_undo(std::vector<std::pair<int, u32>>& log, size_t mark) {
    while (log.size() > mark) { auto [s, v] = log.back(); log.pop_back(); X[s].add(v); }
  }

  // Degree-based candidate pruning, single pass (biplex.cpp:695-713, PRUNE_ONCE).
  void prune_c(std::vector<std::pair<int, u32>>& clog) {
    for (int s = 0; s < 2; ++s) {
      std::vector<u32> snap(C[s].items);
      for (u32 u : snap) {
        if (!C[s].has(u)) continue;
        int64_t d = (int64_t)degC[s][u] + ((int64_t)S[s ^ 1].size() - (int64_t)nonS[s][u].size()) + k;
        if (d < (int64_t)lb) c_del(s, u, clog);
      }
    }
  }

  // Include u (side us) into S: filter C and X (biplex.cpp:1100-1164), then record
  // the new non-neighbour relations (biplex.cpp:1112-1130).
  void include(u32 u, int us, std::vector<std::pair<int, u32>>& clog,
               std::vector<std::pair<int, u32>>& xlog, std::vector<std::pair<int, u32>>& nlog) {
    int os = us ^ 1;
    c_del(us, u, clog);
    x_del(us, u, xlog);
    bool u_full = nonS[us][u].size() >= k;
    for (int pass = 0; pass < 2; ++pass) {
      ISet& V = pass == 0 ? C[os] : X[os];
      std::vector<u32> snap(V.items);
      for (u32 v : snap) {
        if (G->connect(us, u, v)) continue;
        if (u_full || nonS[os][v].size() >= k) {
          if (pass == 0) c_del(os, v, clog); else x_del(os, v, xlog);
        }
      }
    }
    // S members of the other side that u misses and that become saturated force
    // every remaining same-side candidate of u to be adjacent to them.
    for (u32 v : nonS[us][u]) {
      if (nonS[os][v].size() + 1 == k) {
        for (int pass = 0; pass < 2; ++pass) {
          ISet& V = pass == 0 ? C[us] : X[us];
          std::vector<u32> snap(V.items);
          for (u32 w : snap)
            if (!G->connect(os, v, w)) {
              if (pass == 0) c_del(us, w, clog); else x_del(us, w, xlog);
            }
        }
      }
    }
    S[us].add(u);
    for (int pass = 0; pass < 3; ++pass) {
      ISet& V = pass == 0 ? S[os] : pass == 1 ? C[os] : X[os];
      for (u32 v : V.items)
        if (!G->connect(us, u, v)) { nonS[os][v].push_back(u); nlog.push_back({os, v}); }
    }
  }
  void uninclude(u32 u, int us, std::vector<std::pair<int, u32>>& clog, size_t cm,
                 std::vector<std::pair<int, u32>>& xlog, size_t xm,
                 std::vector<std::pair<int, u32>>& nlog, size_t nm) {
    while (nlog.size() > nm) { auto [s, v] = nlog.back(); nlog.pop_back(); nonS[s][v].pop_back(); }
    S[us].del(u);
    x_undo(xlog, xm);
    c_undo(clog, cm);
  }

  // Move u from C to X / back (biplex.cpp:1203-1242).
  void to_x(int s, u32 u, std::vector<std::pair<int, u32>>& clog) { c_del(s, u, clog); X[s].add(u); }

  bool sizes_ok() const {
    return S[0].size() + C[0].size() >= lb && S[1].size() + C[1].size() >= lb;
  }

  void emit() {
    std::vector<u32> L(S[0].items), R;
    for (u32 v : S[1].items) R.push_back(v + G->nL);
    sink->emit(L, R);
    if (sink->full()) stop = true;
  }

  void branch_include(u32 u, int us) {
    std::vector<std::pair<int, u32>> clog, xlog, nlog;
    include(u, us, clog, xlog, nlog);
    if (sizes_ok()) branch();
    uninclude(u, us, clog, 0, xlog, 0, nlog, 0);
  }

  // Main search (biplex.cpp:781-1082).  Upper-bound pruning is omitted: it only
  // removes branches that cannot reach the size bound, so the output is unchanged.
  void branch() {
    if (stop) return;
    ++branches;
    std::vector<std::pair<int, u32>> plog;
    prune_c(plog);
    auto done = [&]() { c_undo(plog, 0); };
    if (!sizes_ok()) { done(); return; }
    if (C[0].size() == 0 && C[1].size() == 0) {
      if (X[0].size() == 0 && X[1].size() == 0 && S[0].size() >= lb && S[1].size() >= lb) emit();
      done(); return;
    }
    const int64_t NEG = -(int64_t)(G->n[0] + G->n[1]) - 1;
    int64_t nd[2] = {NEG, NEG};
    u32 sp[2] = {~0u, ~0u};
    for (int s = 0; s < 2; ++s) {
      for (u32 u : S[s].items) {
        int64_t d = (int64_t)C[s ^ 1].size() - degC[s][u] + (int64_t)nonS[s][u].size();
        if (d > nd[s]) { nd[s] = d; sp[s] = u; }
      }
      if ((int64_t)(S[s ^ 1].size() + C[s ^ 1].size()) - nd[s] < (int64_t)lb - (int64_t)k) { done(); return; }
    }
    if (nd[0] > (int64_t)k || nd[1] > (int64_t)k) {
      // Binary branch on a candidate missed by the S vertex with most misses.
      int side = nd[0] > nd[1] ? 0 : 1;
      int os = side ^ 1;
      u32 u = ~0u;
      int64_t best = G->n[side] + 1, bestNon = -1;
      for (u32 v : C[os].items) {
        int64_t non = nonS[os][v].size();
        int64_t d = (int64_t)degC[os][v] + S[side].size() - non;
        if (((d < best) || (d == best && non > bestNon)) && !G->connect(side, sp[side], v)) {
          u = v; best = d; bestNon = non;
        }
      }
      if (u == ~0u) { done(); return; }
      branch_include(u, os);
      std::vector<std::pair<int, u32>> clog;
      to_x(os, u, clog);
      if (sizes_ok() && !stop) branch();
      X[os].del(u);
      c_undo(clog, 0);
      done(); return;
    }
    // Pivot branching (biplex.cpp:970-1071).
    u32 p[2] = {~0u, ~0u};
    int64_t dp[2] = {-1, -1};
    for (int s = 0; s < 2; ++s) {
      for (u32 u : C[s].items)
        if ((int64_t)degC[s][u] > dp[s]) { dp[s] = degC[s][u]; p[s] = u; }
      for (u32 u : X[s].items) {
        if ((int64_t)degC[s][u] <= dp[s]) continue;
        bool bad = false;
        for (u32 v : nonS[s][u])
          if ((int64_t)C[s].size() - degC[s ^ 1][v] + (int64_t)nonS[s ^ 1][v].size() >= (int64_t)k) { bad = true; break; }
        if (!bad) { dp[s] = degC[s][u]; p[s] = u; }
      }
    }
    int ps = ((int64_t)C[1].size() - dp[0] < (int64_t)C[0].size() - dp[1]) ? 0 : 1;
    u32 pivot = p[ps];
    if (pivot == ~0u) { done(); return; }
    int os = ps ^ 1;
    std::vector<u32> cand;
    for (u32 u : C[os].items)
      if (!G->connect(ps, pivot, u)) cand.push_back(u);
    std::vector<std::pair<int, u32>> mlog;
    for (u32 u : cand) {
      if (stop) break;
      branch_include(u, os);
      to_x(os, u, mlog);
    }
    if (!stop && C[ps].has(pivot)) {
      branch_include(pivot, ps);
    }
    // Backtrack: candidates moved to X come back to C.
    for (auto& e : mlog) X[e.first].del(e.second);
    c_undo(mlog, 0);
    done();
  }
};

// Degeneracy order over the whole bipartite graph (biplex.cpp:85-119): repeatedly
// remove a minimum-degree vertex; ties by smallest global ID.
static void core_order(const BiG& G, std::vector<u32> order[2]) {
  u32 N = G.n[0] + G.n[1];
  std::vector<u32> deg(N);
  u32 maxd = 0;
  for (u32 v = 0; v < N; ++v) {
    deg[v] = v < G.n[0] ? (u32)G.nbr[0][v].size() : (u32)G.nbr[1][v - G.n[0]].size();
    maxd = std::max(maxd, deg[v]);
  }
  std::vector<std::vector<u32>> bucket(maxd + 1);
  for (u32 v = N; v-- > 0;) bucket[deg[v]].push_back(v);  // stacks: pop smallest ID first
  std::vector<char> gone(N, 0);
  u32 cnt[2] = {0, 0};
  order[0].assign(G.n[0], 0); order[1].assign(G.n[1], 0);
  u32 cur = 0;
  for (u32 done = 0; done < N;) {
    while (cur <= maxd && bucket[cur].empty()) ++cur;
    if (cur > maxd) break;
    u32 v = bucket[cur].back(); bucket[cur].pop_back();
    if (gone[v] || deg[v] != cur) continue;
    gone[v] = 1; ++done;
    int s = v < G.n[0] ? 0 : 1;
    u32 lv = s ? v - G.n[0] : v;
    order[s][lv] = cnt[s]++;
    for (u32 w : G.nbr[s][lv]) {
      u32 gw = s ? w : w + G.n[0];
      if (gone[gw] || deg[gw] <= cur) continue;
      --deg[gw];
      bucket[deg[gw]].push_back(gw);
    }
    if (cur > 0) --cur;
  }
}

// Seed sets for seeded mode (biplex.cpp:227-393): C[opp] = N(u), C[same] = later
// 2-hop vertices, X[same] = earlier 2-hop vertices, degree pruning, then 3-hop
// opposite-side vertices with enough candidate neighbours.
static bool seed_prune(BPP& B, u32 u, int us, u32 q) {
  int os = us ^ 1;
  u32 k = B.k;
  std::vector<std::pair<int, u32>> dummy;
  for (;;) {
    bool changed = false;
    std::vector<u32> snap(B.C[os].items);
    for (u32 v : snap) {
      int64_t d = (int64_t)B.degC[os][v] + (B.G->connect(us, u, v) ? 1 : 0);
      if (d < (int64_t)q - (int64_t)k) {
        changed = true;
        B.c_del(os, v, dummy);
        if ((int64_t)B.degC[us][u] < (int64_t)q - (int64_t)k) return false;
      }
    }
    if (changed) continue;
    snap = B.C[us].items;
    for (u32 v : snap)
      if ((int64_t)B.degC[us][v] < (int64_t)q - 2 * (int64_t)k) { changed = true; B.c_del(us, v, dummy); }
    snap = B.X[us].items;
    for (u32 v : snap)
      if ((int64_t)B.degC[us][v] < (int64_t)q - 2 * (int64_t)k) { changed = true; B.X[us].del(v); }
    if (!changed) break;
  }
  return true;
}

static void run_seeded(BPP& B, const BiG& G, u32 q) {
  std::vector<u32> order[2];
  core_order(G, order);
  u32 maxd[2] = {0, 0};
  for (int s = 0; s < 2; ++s)
    for (auto& l : G.nbr[s]) maxd[s] = std::max(maxd[s], (u32)l.size());
  int us = maxd[0] < maxd[1] ? 0 : 1;
  int os = us ^ 1;
  std::vector<u32> ordered(G.n[us]);
  for (u32 v = 0; v < G.n[us]; ++v) ordered[order[us][v]] = v;
  for (u32 i = 0; i < G.n[us] && !B.stop; ++i) {
    u32 u = ordered[i];
    B.init_sets();
    B.S[us].add(u);
    for (u32 v : G.nbr[us][u]) {
      B.C[os].add(v);
      for (u32 w : G.nbr[os][v]) {
        if (w == u) continue;
        if (order[us][w] > order[us][u]) B.C[us].add(w); else B.X[us].add(w);
      }
    }
    for (int s = 0; s < 2; ++s)
      for (u32 v : B.C[s].items)
        for (u32 w : G.nbr[s][v]) ++B.degC[s ^ 1][w];
    if (!seed_prune(B, u, us, q)) continue;
    std::vector<u32> snap(B.C[us].items);
    for (u32 v : snap)
      for (u32 w : G.nbr[us][v]) {
        if (B.C[os].has(w)) continue;
        if ((int64_t)B.degC[os][w] + (int64_t)B.k >= (int64_t)q) {
          B.C[os].add(w);
          for (u32 x : G.nbr[os][w]) ++B.degC[us][x];
        }
      }
    if (!seed_prune(B, u, us, q)) continue;
    for (int pass = 0; pass < 2; ++pass) {
      ISet& V = pass == 0 ? B.C[os] : B.X[os];
      for (u32 v : V.items)
        if (!G.connect(us, u, v)) B.nonS[os][v].push_back(u);
    }
    B.branch();
  }
}

extern "C" {

// Enumerate maximal k-biplexes with |L| >= theta and |R| >= theta.
// mode 0: global search (biplex.cpp:480-497); mode 1: seeded search
// (biplex.cpp:498-549; valid only when theta > 2k); mode 2: seeded when
// theta > 2k else global, as the reference chooses.  Stops after `limit` results.
// Returns the number of results; *out_words receives malloc'd records (caller
// frees with or_free) when keep != 0.
long long or_bppivot(u32 nLeft, u32 N, const u32* off, const u32* adj, u32 k, u32 theta,
                     u64 limit, int mode, int keep, u32** out_words, u64* out_nwords,
                     or_digest* dg, u64* branches) {
  BiG G;
  G.build(nLeft, N, off, adj);
  Sink sink;
  sink.keep = keep != 0;
  sink.limit = limit ? limit : ~0ull;
  BPP B;
  B.G = &G; B.k = k; B.lb = theta; B.sink = &sink;
  B.init_sets();
  bool seeded = mode == 1 || (mode == 2 && theta > 2 * k);
  if (seeded) {
    run_seeded(B, G, theta);
  } else {
    for (int s = 0; s < 2; ++s) {
      B.degC[s].assign(G.n[s], 0);
      for (u32 v = 0; v < G.n[s]; ++v) {
        B.C[s].add(v);
        B.degC[s][v] = (u32)G.nbr[s][v].size();
      }
    }
    B.branch();
  }
  if (branches) *branches = B.branches;
  if (dg) { dg->count = sink.count; dg->fp_sum = sink.fp_sum; dg->fp_xor = sink.fp_xor; dg->canon_lo = dg->canon_hi = 0; }
  if (out_words && out_nwords) {
    *out_nwords = sink.words.size();
    *out_words = (u32*)malloc(sizeof(u32) * (sink.words.size() + 1));
    if (!sink.words.empty()) memcpy(*out_words, sink.words.data(), sizeof(u32) * sink.words.size());
  }
  return (long long)sink.count;
}

void or_free(void* p) { free(p); }

// Exhaustive enumeration over all (L, R) pairs: |U| + |V| <= 24.
long long or_brute(u32 nLeft, u32 N, const u32* off, const u32* adj, u32 k, u32 theta,
                   u32** out_words, u64* out_nwords, or_digest* dg) {
  u32 nR = N - nLeft;
  if (nLeft > 24 || nR > 24 || nLeft + nR > 24) return -1;
  std::vector<u32> adjL(nLeft, 0), adjR(nR, 0);  // bitmasks
  for (u32 u = 0; u < nLeft; ++u)
    for (u32 e = off[u]; e < off[u + 1]; ++e) adjL[u] |= 1u << (adj[e] - nLeft);
  for (u32 v = 0; v < nR; ++v)
    for (u32 e = off[nLeft + v]; e < off[nLeft + v + 1]; ++e) adjR[v] |= 1u << adj[e];
  auto ok = [&](u32 Lm, u32 Rm) {
    for (u32 u = 0; u < nLeft; ++u)
      if ((Lm >> u) & 1)
        if ((u32)__builtin_popcount(Rm & ~adjL[u]) > k) return false;
    for (u32 v = 0; v < nR; ++v)
      if ((Rm >> v) & 1)
        if ((u32)__builtin_popcount(Lm & ~adjR[v]) > k) return false;
    return true;
  };
  Sink sink;
  for (u32 Lm = 1; Lm < (1u << nLeft); ++Lm) {
    if ((u32)__builtin_popcount(Lm) < theta) continue;
    for (u32 Rm = 1; Rm < (1u << nR); ++Rm) {
      if ((u32)__builtin_popcount(Rm) < theta) continue;
      if (!ok(Lm, Rm)) continue;
      bool maximal = true;
      for (u32 w = 0; w < nLeft && maximal; ++w)
        if (!((Lm >> w) & 1) && ok(Lm | (1u << w), Rm)) maximal = false;
      for (u32 y = 0; y < nR && maximal; ++y)
        if (!((Rm >> y) & 1) && ok(Lm, Rm | (1u << y))) maximal = false;
      if (!maximal) continue;
      std::vector<u32> L, R;
      for (u32 u = 0; u < nLeft; ++u) if ((Lm >> u) & 1) L.push_back(u);
      for (u32 v = 0; v < nR; ++v) if ((Rm >> v) & 1) R.push_back(v + nLeft);
      sink.emit(L, R);
    }
  }
  if (dg) { dg->count = sink.count; dg->fp_sum = sink.fp_sum; dg->fp_xor = sink.fp_xor; dg->canon_lo = dg->canon_hi = 0; }
  if (out_words && out_nwords) {
    *out_nwords = sink.words.size();
    *out_words = (u32*)malloc(sizeof(u32) * (sink.words.size() + 1));
    if (!sink.words.empty()) memcpy(*out_words, sink.words.data(), sizeof(u32) * sink.words.size());
  }
  return (long long)sink.count;
}

// Digest of a record stream.  canon: records sorted lexicographically as u32
// sequences, then hashed word by word with two independent 64-bit lanes.
// fp_sum/fp_xor: the order-free per-record fingerprints.
int or_canon_digest(const u32* words, u64 nwords, or_digest* dg) {
  std::vector<std::pair<u64, u64>> recs;  // (offset, length)
  u64 i = 0;
  dg->count = dg->fp_sum = dg->fp_xor = 0;
  while (i < nwords) {
    u64 s = i;
    u32 nl = words[i];
    if (i + 1 + nl >= nwords) return -1;
    u32 nr = words[i + 1 + nl];
    u64 len = 2 + (u64)nl + nr;
    if (s + len > nwords) return -1;
    u64 f = fp_of(words + s + 1, nl, words + s + 2 + nl, nr);
    dg->count++; dg->fp_sum += f; dg->fp_xor ^= f;
    recs.push_back({s, len});
    i = s + len;
  }
  std::sort(recs.begin(), recs.end(), [&](const std::pair<u64, u64>& a, const std::pair<u64, u64>& b) {
    return std::lexicographical_compare(words + a.first, words + a.first + a.second,
                                        words + b.first, words + b.first + b.second);
  });
  u64 h1 = 0x6A09E667F3BCC908ull, h2 = 0xBB67AE8584CAA73Bull;
  for (auto& r : recs)
    for (u64 j = 0; j < r.second; ++j) {
      u64 w = words[r.first + j];
      h1 = or_mix(h1 ^ (w + 0x100000000ull));
      h2 = or_mix(h2 + (w * 0xD6E8FEB86659FD93ull) + 0x3C6EF372FE94F82Bull);
    }
  dg->canon_lo = h1; dg->canon_hi = h2 ^ (u64)recs.size();
  return 0;
}

// Validate a record stream against G: every record is a k-biplex with both sides
// >= theta, maximal in G, and no record repeats.  Returns 0 when valid, else the
// 1-based index of the first bad record; *why gets 1 = malformed, 2 = not a
// k-biplex, 3 = below theta, 4 = not maximal, 5 = duplicate, 6 = wrong side.
long long or_validate(u32 nLeft, u32 N, const u32* off, const u32* adj, u32 k, u32 theta,
                      const u32* words, u64 nwords, int* why) {
  std::vector<u32> cnt(N, 0), inset(N, 0);
  std::vector<u64> fps;
  u64 i = 0, idx = 0;
  u32 stamp = 0;
  *why = 0;
  while (i < nwords) {
    ++idx; ++stamp;
    u32 nl = words[i];
    if (i + 1 + nl >= nwords) { *why = 1; return (long long)idx; }
    const u32* L = words + i + 1;
    u32 nr = words[i + 1 + nl];
    const u32* R = words + i + 2 + nl;
    if (i + 2 + nl + nr > nwords) { *why = 1; return (long long)idx; }
    i += 2 + nl + nr;
    if (nl < theta || nr < theta) { *why = 3; return (long long)idx; }
    for (u32 a = 0; a < nl; ++a) if (L[a] >= nLeft) { *why = 6; return (long long)idx; }
    for (u32 a = 0; a < nr; ++a) if (R[a] < nLeft || R[a] >= N) { *why = 6; return (long long)idx; }
    for (u32 a = 0; a < nl; ++a) inset[L[a]] = stamp;
    for (u32 a = 0; a < nr; ++a) inset[R[a]] = stamp;
    // cnt[x] = neighbours of x inside the opposite side of the record.
    std::vector<u32> touched;
    auto scan = [&](const u32* A, u32 na) {
      for (u32 a = 0; a < na; ++a)
        for (u32 e = off[A[a]]; e < off[A[a] + 1]; ++e) {
          u32 w = adj[e];
          if (cnt[w]++ == 0) touched.push_back(w);
        }
    };
    scan(L, nl); scan(R, nr);
    auto miss = [&](u32 x, u32 other) { return other - cnt[x]; };
    bool bad = false;
    for (u32 a = 0; a < nl && !bad; ++a) if (miss(L[a], nr) > k) bad = true;
    for (u32 a = 0; a < nr && !bad; ++a) if (miss(R[a], nl) > k) bad = true;
    if (bad) { *why = 2; return (long long)idx; }
    // Maximality: w outside can be added iff w misses <= k of the other side and
    // every vertex it misses there has fewer than k misses already.
    auto addable = [&](u32 w, bool left) -> bool {
      const u32* O = left ? R : L;
      u32 no = left ? nr : nl;
      if (no - cnt[w] > k) return false;
      // vertices of O not adjacent to w
      u32 j = off[w], je = off[w + 1];
      for (u32 a = 0; a < no; ++a) {
        while (j < je && adj[j] < O[a]) ++j;
        bool adjw = j < je && adj[j] == O[a];
        if (!adjw) {
          u32 other = left ? nl : nr;
          if (miss(O[a], other) >= k) return false;
        }
      }
      return true;
    };
    // Candidates adjacent to the record are checked one by one.  A vertex with no
    // neighbour in the other side misses all of it: it can be added only when that
    // side has at most k vertices, none of them already at k misses; all such
    // vertices behave alike, so one existence test covers them.
    u32 touchedL = 0, touchedR = 0;
    for (u32 w : touched) {
      if (inset[w] == stamp) continue;
      if (w < nLeft) { ++touchedL; if (!bad && addable(w, true)) bad = true; }
      else { ++touchedR; if (!bad && addable(w, false)) bad = true; }
    }
    if (!bad && nr <= k) {
      bool sat = false;
      for (u32 a = 0; a < nr; ++a) if (miss(R[a], nl) >= k) sat = true;
      if (!sat && (u64)nLeft - nl - touchedL > 0) bad = true;
    }
    if (!bad && nl <= k) {
      bool sat = false;
      for (u32 a = 0; a < nl; ++a) if (miss(L[a], nr) >= k) sat = true;
      if (!sat && (u64)(N - nLeft) - nr - touchedR > 0) bad = true;
    }
    for (u32 w : touched) cnt[w] = 0;
    if (bad) { *why = 4; return (long long)idx; }
    fps.push_back(fp_of(L, nl, R, nr));
  }
  // Duplicates: equal fingerprints are compared record by record is unnecessary
  // here because fingerprints of distinct sets collide with probability ~2^-64
  // per pair; an equal pair is reported as a duplicate.
  std::vector<u64> s(fps);
  std::sort(s.begin(), s.end());
  for (size_t a = 1; a < s.size(); ++a)
    if (s[a] == s[a - 1]) { *why = 5; return (long long)idx; }
  return 0;
}

u64 or_fingerprint(const u32* L, u32 nl, const u32* R, u32 nr) { return fp_of(L, nl, R, nr); }

// Fingerprint of every record of a stream into out (capacity cap); returns the
// record count, or -1 when the stream is malformed or out is too small.
long long or_fingerprints(const u32* words, u64 nwords, u64* out, u64 cap) {
  u64 i = 0, n = 0;
  while (i < nwords) {
    const u32 nl = words[i];
    if (i + 1 + nl >= nwords) return -1;
    const u32 nr = words[i + 1 + nl];
    if (i + 2 + nl + nr > nwords || n >= cap) return -1;
    out[n++] = fp_of(words + i + 1, nl, words + i + 2 + nl, nr);
    i += 2 + nl + nr;
  }
  return (long long)n;
}

// Offsets (in words) of the records of a stream; returns the record count or -1.
long long or_record_offsets(const u32* words, u64 nwords, u64* out, u64 cap) {
  u64 i = 0, n = 0;
  while (i < nwords) {
    if (n >= cap) return -1;
    out[n++] = i;
    const u32 nl = words[i];
    if (i + 1 + nl >= nwords) return -1;
    const u32 nr = words[i + 1 + nl];
    i += 2 + nl + nr;
  }
  return i == nwords ? (long long)n : -1;
}

}  // extern "C"
