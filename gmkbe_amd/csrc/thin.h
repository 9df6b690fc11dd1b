// Thin results: maximal k-biplexes with one side ("side A", the tuple side) of
// at most 2k vertices.  Such a result cannot be found inside a 2-hop seed
// partition: members of the small side need not share neighbours, and with at
// most k vertices on a side every vertex of the other side may be a member.
//
// For a tuple L = (l_0 < ... < l_{j-1}) on side A (j <= 2k) the other side R of
// every maximal result (L, R) is built from miss types.  For a vertex y of side
// B its type is the set of tuple members it is NOT adjacent to (a j-bit mask).
// Types with more than k bits are infeasible; type 0 vertices ("free") are in
// every maximal R; vertices of a type m are interchangeable for L, so R is fixed
// by a count c_m per type plus which c_m vertices of each pool are taken.  Counts
// obey one budget per member (sum of c_m over masks containing member i <= k) and
// are maximal (a type with spare vertices has a member whose budget is spent).
// Vertices adjacent to no member form the far pool (type = all bits, only
// feasible when j <= k); it is never materialised: member i of the far pool is
// found by select over the sorted list N(L).
//
// A concrete R is reported iff (L, R) is maximal on side A too: no w outside L
// can be added.  w can be added iff it misses at most k of R and is adjacent to
// every saturated y in R (popcount(type) == k).  Candidates are the neighbours
// of one saturated y, or, without saturated vertices and |R| > k, the
// neighbours of the k+1 lowest-degree members of R (w misses at most k of them).
//
// This is the role of GMkBPE's Generate/Special path (include/for_generate.cuh:
// 300-512, include/for_special.cuh:274-400: conY = free vertices, varY = pools,
// k-combinations checked by CheckMaximality_for_special,
// include/for_checkmaximality.cuh:2587-2744), generalised to |L| <= 2k.
//
// All functions here are plain C++ usable on the host and, compiled by hipcc,
// on the device (one lane per tuple / per combination).
#pragma once
#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define GMK_HD __host__ __device__ __forceinline__
#else
#define GMK_HD inline
#endif

namespace gmk {
namespace thin {

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;

constexpr int MAXJ = 6;       // tuple size limit (2k with k <= 3)
constexpr int MAXT = 18;      // chosen types per family (<= j * k)
constexpr u64 SAT = 1ull << 62;  // saturation value of combination counts
constexpr u64 FARP = ~0ull;       // pool address marking the far pool

GMK_HD int popc(u32 x) {
  int c = 0;
  while (x) { x &= x - 1; ++c; }
  return c;
}

// C(n, c) for c <= 3, saturating at SAT.
GMK_HD u64 binom(u64 n, u32 c) {
  if (c == 0) return 1;
  if (n < c) return 0;
  if (c == 1) return n;
  if (c == 2) {
    const u64 a = n % 2 == 0 ? n / 2 : n, b = n % 2 == 0 ? n - 1 : (n - 1) / 2;
    if (a && b > SAT / a) return SAT;
    return a * b;
  }
  // c == 3: n (n-1) (n-2) / 6
  u64 x = n, y = n - 1, z = n - 2;
  if (x % 2 == 0) x /= 2; else y /= 2;
  if (x % 3 == 0) x /= 3; else if (y % 3 == 0) y /= 3; else z /= 3;
  if (x && y > SAT / x) return SAT;
  const u64 xy = x * y;
  if (xy && z > SAT / xy) return SAT;
  return xy * z;
}

GMK_HD u64 mul_sat(u64 a, u64 b) {
  if (a == 0 || b == 0) return 0;
  if (a >= SAT || b >= SAT || a > SAT / b) return SAT;
  return a * b;
}

// min(a + b, SAT) for a, b <= SAT (associative, so any scan order gives the
// same prefix sums).
GMK_HD u64 add_sat(u64 a, u64 b) {
  const u64 z = a + b;
  return (z < a || z > SAT) ? SAT : z;
}

// Exclusive prefix sums of family totals, saturating at SAT, as computed by one
// block of TSCAN threads (thin_scan_totals): thread t owns a contiguous
// segment of families, the segment sums are scanned across the block, and each
// thread writes its segment's prefixes from the sum of the segments before it.
// The phases are shared with the CPU model that tests them.
constexpr unsigned TSCAN = 1024;
GMK_HD void scan_segment(u32 n, u32 t, u32& b, u32& e) {
  const u32 seg = (n + TSCAN - 1) / TSCAN;
  b = (u64)t * seg < n ? t * seg : n;
  e = b + seg < n ? b + seg : n;
}
// Phase 1: the sum of thread t's segment.
template <class TotalAt>
GMK_HD u64 scan_phase_sum(u32 n, u32 t, TotalAt total) {
  u32 b, e;
  scan_segment(n, t, b, e);
  u64 s = 0;
  for (u32 i = b; i < e; ++i) s = add_sat(s, total(i));
  return s;
}
// Phase 2 (one step of an inclusive Hillis-Steele scan over the segment sums,
// d = 1, 2, 4, ..., TSCAN / 2): thread t's new value from the old values.
GMK_HD u64 scan_phase_step(const u64* part, u32 t, u32 d) { return t >= d ? add_sat(part[t], part[t - d]) : part[t]; }
// Phase 3: thread t writes its segment's exclusive prefixes, given the
// inclusive scan `incl` of the segment sums; the last thread writes the total.
template <class TotalAt>
GMK_HD void scan_phase_write(u32 n, u32 t, const u64* incl, TotalAt total, u64* prefix) {
  u32 b, e;
  scan_segment(n, t, b, e);
  u64 run = t ? incl[t - 1] : 0;
  for (u32 i = b; i < e; ++i) { prefix[i] = run; run = add_sat(run, total(i)); }
  if (t == TSCAN - 1) prefix[n] = incl[TSCAN - 1];
}

// Colex unrank: rank r of C(n, c) -> idx[0] < ... < idx[c-1].
GMK_HD void unrank(u64 r, u32 c, u64 n, u32* idx) {
  // c <= 3; constant indices keep idx in registers
#pragma unroll
  for (int t = 3; t >= 1; --t) {
    if ((u32)t > c) continue;
    // largest x in [t-1, n-1] with C(x, t) <= r
    u64 lo = t - 1, hi = n - 1;
    while (lo < hi) {
      const u64 mid = lo + (hi - lo + 1) / 2;
      if (binom(mid, t) <= r) lo = mid; else hi = mid - 1;
    }
    idx[t - 1] = (u32)lo;
    r -= binom(lo, t);
    n = lo;
  }
}

// Algorithmic byte count of the memory the functions below touch: every array
// element read or written counts its size (binary-search probes included).
struct Cost {
  u64 bytes = 0;
};

// Sorted-list membership.
GMK_HD bool contains(const u32* a, u32 n, u32 x, Cost& cost) {
  u32 lo = 0, hi = n;
  u32 probes = 0;
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    ++probes;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  const bool hit = lo < n && a[lo] == x;
  cost.bytes += 4ull * (probes + (lo < n ? 1 : 0));
  return hit;
}

// i-th vertex of [b_lo, b_hi) that is not in the sorted list nl (size d).
GMK_HD u32 far_select(const u32* nl, u32 d, u32 b_lo, u64 i, Cost& cost) {
  // p = number of members m_q with (m_q - b_lo - q) <= i
  u32 lo = 0, hi = d;
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    cost.bytes += 4;
    if ((u64)(nl[mid] - b_lo) - mid <= i) lo = mid + 1; else hi = mid;
  }
  return (u32)(b_lo + i + lo);
}

// A family: one tuple with one maximal count vector.  Lists are addressed by
// pointer (arena or, for one-vertex tuples, the CSR itself).
struct Family {
  u32 L[MAXJ];
  u32 j;
  u32 nfree;          // free vertices (type 0)
  u32 ntypes;         // types with c > 0
  u32 rsize;          // |R| = nfree + sum c
  const u32* freep;   // free list, ascending
  const u32* nlp;     // N(L) ascending (far-pool select); valid when a far pool is used
  u32 nl_size;
  u32 tuple;          // index of the tuple in its batch
  u32 mask[MAXT];
  u32 cnt[MAXT];
  u32 psize[MAXT];
  u64 pool[MAXT];     // address of the pool, FARP for the far pool
  u64 total;          // number of concrete R (product of binomials), saturating
  u32 fpick[4];       // min(k + 1, nfree) free vertices of least degree
  u32 nfpick;
};

// The min(k + 1, nfree) free vertices of least degree (ties: first in the list).
GMK_HD u32 free_picks(const u32* off, const u32* freel, u32 nfree, u32 k, u32* pick, Cost& cost) {
  u32 pdeg[4];
  u32 np = 0;
  for (u32 i = 0; i < nfree; ++i) {
    const u32 y = freel[i], d = off[y + 1] - off[y];
    if (np < k + 1) { pick[np] = y; pdeg[np] = d; ++np; continue; }
    u32 worst = 0;
    for (u32 q = 1; q < np; ++q) if (pdeg[q] > pdeg[worst]) worst = q;
    if (d < pdeg[worst]) { pick[worst] = y; pdeg[worst] = d; }
  }
  cost.bytes += 12ull * nfree;
  return np;
}

// The y of index i in pool t of family f.
GMK_HD u32 pool_member(const Family& f, u32 t, u64 i, u32 b_lo, Cost& cost) {
  cost.bytes += 4;
  if (f.pool[t] == FARP) return far_select(f.nlp, f.nl_size, b_lo, i, cost);
  return ((const u32*)f.pool[t])[i];
}

// Neighbour hashes: every vertex v of degree d >= nh_min has an open-addressing
// set of its neighbours, 2^nh_log2(d) >= 2d slots (load <= 1/2) from
// nh[nh_off[v]], empty slots NH_EMPTY, linear probing from nh_slot.  A
// membership test then reads about 1.5 slots instead of log2(d) binary-search
// probes of a long list.
constexpr u32 NH_EMPTY = 0xFFFFFFFFu;
GMK_HD u32 nh_log2(u32 d) { return 32u - (u32)__builtin_clz(2u * d - 1u); }
GMK_HD u32 nh_slot(u32 x, u32 lg) { return (u32)(((u64)(x * 0x9E3779B1u) << lg) >> 32); }

struct Graph {
  const u32* off;
  const u32* adj;
  u32 a_lo, a_hi;  // side A (tuple side), global IDs
  u32 b_lo, b_hi;  // side B
  const u64* nh_off = nullptr;  // neighbour hashes (none when nh is null)
  const u32* nh = nullptr;
  u32 nh_min = 0xFFFFFFFFu;
};

// The neighbour hash of v (degree d), or null when v has none.
GMK_HD const u32* nh_table(const Graph& g, u32 v, u32 d, Cost& cost) {
  if (!g.nh || d < g.nh_min) return nullptr;
  cost.bytes += 8;
  return g.nh + g.nh_off[v];
}

// Is x in the neighbour hash t of a vertex of degree d?
GMK_HD bool nh_find(const u32* t, u32 d, u32 x, Cost& cost) {
  const u32 lg = nh_log2(d), mask = (1u << lg) - 1u;  // lg <= 31: d < 2^30
  for (u32 s = nh_slot(x, lg);; s = (s + 1) & mask) {
    const u32 y = t[s];
    cost.bytes += 4;
    if (y == x) return true;
    if (y == NH_EMPTY) return false;
  }
}

// Is x a neighbour of a vertex with sorted list nv of size d and hash t
// (null: none)?
GMK_HD bool adjacent(const u32* t, const u32* nv, u32 d, u32 x, Cost& cost) {
  return t ? nh_find(t, d, x, cost) : contains(nv, d, x, cost);
}

// Threat test: can some w in side A outside L be added to (L, R)?
//   R = free (sorted, nfree) + ch (nch chosen, each with its type size tsz).
GMK_HD bool in_tuple(const u32* L, u32 j, u32 w) {
  for (u32 i = 0; i < j; ++i)
    if (L[i] == w) return true;
  return false;
}

// Does w miss at most k vertices of R, and is w adjacent to every saturated one?
// `known` is a member of R already known to be adjacent to w (the vertex whose
// list produced w as a candidate), or NONE.  Members are looked up in N(w)
// (its hash, when w has one).
// NMAX bounds nch at compile time so the chosen vertices stay in registers.
template <int NMAX>
GMK_HD bool addable(const Graph& g, u32 w, const u32* freel, u32 nfree, const u32* ch, const u32* ctsz, u32 nch,
                    u32 k, u32 known, Cost& cost) {
  const u32 ow = g.off[w], dw = g.off[w + 1] - ow;
  const u32* nw = g.adj + ow;
  cost.bytes += 8;
  const u32 rs = nfree + nch;
  if (dw + k < rs) return false;
  const u32* hw = nh_table(g, w, dw, cost);
  u32 miss = 0;
  bool out = false;  // a saturated member is missed, or more than k members
#pragma unroll
  for (int i = 0; i < NMAX; ++i) {
    if ((u32)i < nch && !out && ch[i] != known && !adjacent(hw, nw, dw, ch[i], cost))
      out = ctsz[i] >= k || ++miss > k;
  }
  if (out) return false;
  // free vertices have type 0 (< k for k >= 1): only their number matters
  if (nfree <= dw) {
    for (u32 i = 0; i < nfree; ++i)
      if (!adjacent(hw, nw, dw, freel[i], cost) && ++miss > k) return false;
    cost.bytes += 4ull * nfree;
  } else {
    // count free vertices adjacent to w by scanning N(w); stop once the rest of
    // N(w) cannot bring the misses back to k
    const u32 need = nfree - (k - miss);  // hits required (miss <= k here)
    u32 hit = 0;
    for (u32 i = 0; i < dw; ++i) {
      if (hit + (dw - i) < need) { cost.bytes += 4ull * i; return false; }
      hit += contains(freel, nfree, nw[i], cost) ? 1 : 0;
    }
    cost.bytes += 4ull * dw;
    miss += nfree - hit;
    if (miss > k) return false;
  }
  return true;
}

// Candidates for a vertex that could join (L, R): up to k + 1 lists of
// vertices of side A (neighbour lists), each with a member already known to be
// adjacent to its entries, and optionally a vertex every candidate must be
// adjacent to.  `decided` >= 0 when no candidate is needed (1 = maximal).
struct CandSegs {
  const u32* lst[4];
  u32 len[4];
  u32 known[4];
  u32 n;
  u32 must;          // candidates must be adjacent to this vertex (NONE: no filter)
  const u32* must_nb;
  const u32* must_nh;  // its neighbour hash, or null
  u32 must_deg;
  int decided;
};

// The candidate lists of (L, R).  fpick holds min(k + 1, nfree) free vertices
// (free_picks).
//  - With a saturated member: an addable w is adjacent to the saturated member
//    `best` of least degree and to all but k members of R, so it neighbours one
//    of any k + 1 other members.  When the k + 1 other members of least degree
//    have fewer neighbours in all than `best`, their lists give the candidates
//    (kept if adjacent to `best`); otherwise N(best) does.
//  - With none: an addable w misses at most k members of R, so it is adjacent
//    to one of any k + 1 members; the k + 1 members of least degree among the
//    chosen vertices and fpick give the candidates.
template <int NMAX>
GMK_HD void candidate_segments(const Graph& g, u32 j, u32 nfree, const u32* fpick, u32 nfpick, const u32* ch,
                               const u32* ctsz, u32 nch, u32 k, CandSegs& cs, Cost& cost) {
  cs.n = 0;
  cs.must = 0xFFFFFFFFu;
  cs.must_nb = nullptr;
  cs.must_nh = nullptr;
  cs.must_deg = 0;
  cs.decided = -1;
  const u32 rs = nfree + nch;
  // saturated chosen vertex with the smallest degree
  u32 best = 0xFFFFFFFFu, bestdeg = 0xFFFFFFFFu;
#pragma unroll
  for (int i = 0; i < NMAX; ++i) {
    if ((u32)i >= nch) break;
    if (ctsz[i] >= k) {
      const u32 d = g.off[ch[i] + 1] - g.off[ch[i]];
      cost.bytes += 8;
      if (d < bestdeg) { bestdeg = d; best = ch[i]; }
    }
  }
  if (best == 0xFFFFFFFFu && rs <= k) {
    // every vertex of side A outside L can be added
    cs.decided = (u64)(g.a_hi - g.a_lo) <= j ? 1 : 0;
    return;
  }
  // the k + 1 members of least degree (excluding `best`); k + 1 <= 4 slots
  // addressed by constant indices so they stay in registers
  u32 pk[4] = {0, 0, 0, 0}, pd[4] = {0, 0, 0, 0};
  u32 np = 0;
  auto take = [&](u32 y) {
    const u32 d = g.off[y + 1] - g.off[y];
    cost.bytes += 8;
    if (np < k + 1) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if ((u32)q == np) { pk[q] = y; pd[q] = d; }
      ++np;
      return;
    }
    u32 worst = 0, wd = pd[0];
#pragma unroll
    for (int q = 1; q < 4; ++q)
      if ((u32)q < np && pd[q] > wd) { worst = q; wd = pd[q]; }
    if (d < wd) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if ((u32)q == worst) { pk[q] = y; pd[q] = d; }
    }
  };
#pragma unroll
  for (int i = 0; i < NMAX; ++i)
    if ((u32)i < nch && ch[i] != best) take(ch[i]);
  for (u32 i = 0; i < nfpick; ++i) take(fpick[i]);
  u64 psum = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) psum += (u32)q < np ? pd[q] : 0u;
  if (best != 0xFFFFFFFFu && !(np == k + 1 && psum < bestdeg)) {
    cs.lst[0] = g.adj + g.off[best];
    cs.len[0] = bestdeg;
    cs.known[0] = best;
    cs.n = 1;
    return;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bool on = (u32)q < np;
    cs.lst[q] = on ? g.adj + g.off[pk[q]] : g.adj;
    cs.len[q] = on ? pd[q] : 0u;
    cs.known[q] = on ? pk[q] : 0xFFFFFFFFu;
  }
  cs.n = np;
  if (best != 0xFFFFFFFFu) {
    cs.must = best;
    cs.must_nb = g.adj + g.off[best];
    cs.must_nh = nh_table(g, best, bestdeg, cost);
    cs.must_deg = bestdeg;
  }
}

// Can candidate w be added?  `known`: the member whose list produced w.
template <int NMAX>
GMK_HD bool candidate_joins(const Graph& g, const u32* L, u32 j, const u32* freel, u32 nfree, const u32* ch,
                            const u32* ctsz, u32 nch, u32 k, const CandSegs& cs, u32 known, u32 w, Cost& cost) {
  if (in_tuple(L, j, w)) return false;
  if (cs.must != 0xFFFFFFFFu && !adjacent(cs.must_nh, cs.must_nb, cs.must_deg, w, cost)) return false;
  return addable<NMAX>(g, w, freel, nfree, ch, ctsz, nch, k, known, cost);
}

// 1 if (L, R) is maximal on side A, 0 if not; -1 when the first `budget`
// candidates (in list order) cannot join and more remain: the caller tests the
// rest another way.
template <int NMAX>
GMK_HD int side_a_maximal(const Graph& g, const u32* L, u32 j, const u32* freel, u32 nfree, const u32* fpick,
                          u32 nfpick, const u32* ch, const u32* ctsz, u32 nch, u32 k, Cost& cost,
                          u32 budget = 0xFFFFFFFFu) {
  CandSegs cs;
  candidate_segments<NMAX>(g, j, nfree, fpick, nfpick, ch, ctsz, nch, k, cs, cost);
  if (cs.decided >= 0) return cs.decided;
  u32 tested = 0;
  // at most 4 lists, visited by constant index (cs stays in registers)
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if ((u32)q >= cs.n) break;
    const u32* lst = cs.lst[q];
    const u32 len = cs.len[q], known = cs.known[q];
    for (u32 e = 0; e < len; ++e) {
      if (tested++ == budget) return -1;
      cost.bytes += 4;
      if (candidate_joins<NMAX>(g, L, j, freel, nfree, ch, ctsz, nch, k, cs, known, lst[e], cost)) return 0;
    }
  }
  return 1;
}

}  // namespace thin
}  // namespace gmk
