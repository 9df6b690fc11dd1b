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
#include <cmath>
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

// a * b, or SAT when it reaches SAT (no division: the product is checked for
// 64-bit overflow, then compared).
GMK_HD u64 mul_sat(u64 a, u64 b) {
  if (a == 0 || b == 0) return 0;
  if (a >= SAT || b >= SAT) return SAT;
  u64 p;
  if (__builtin_mul_overflow(a, b, &p)) return SAT;
  return p < SAT ? p : SAT;
}

// C(n, c) for c <= 3, saturating at SAT.
GMK_HD u64 binom(u64 n, u32 c) {
  if (c == 0) return 1;
  if (n < c) return 0;
  if (c == 1) return n < SAT ? n : SAT;
  if (n >= (1ull << 32)) return SAT;  // pools are smaller; keeps n (n - 1) below 2^64
  const u64 p2 = (n * (n - 1)) >> 1;   // C(n, 2) < 2^63
  if (c == 2) return p2 < SAT ? p2 : SAT;
  // c == 3: C(n, 2) (n - 2) / 3, with the factor 3 taken out of whichever term has it
  if (p2 % 3 == 0) return mul_sat(p2 / 3, n - 2);
  return mul_sat(p2, (n - 2) / 3);
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

// Largest x in [t - 1, n - 1] with C(x, t) <= r (t <= 3, r < C(n, t)): a
// floating-point estimate from the inverse of C(x, t) ~ x^t / t!, then exact
// integer steps to the answer (the estimate is off by at most a few).
GMK_HD u64 colex_top(u64 r, u32 t, u64 n) {
  if (t == 1) return r;  // C(x, 1) = x
  const double rd = (double)r;
  double e = t == 2 ? 0.5 * (1.0 + sqrt(8.0 * rd + 1.0)) : cbrt(6.0 * rd) + 1.0;
  const double lo = (double)(t - 1), hi = (double)(n - 1);
  e = e < lo ? lo : (e > hi ? hi : e);
  u64 x = (u64)e;
  while (x > t - 1 && binom(x, t) > r) --x;
  while (x + 1 <= n - 1 && binom(x + 1, t) <= r) ++x;
  return x;
}

// Colex unrank: rank r of C(n, c) -> idx[0] < ... < idx[c-1].
GMK_HD void unrank(u64 r, u32 c, u64 n, u32* idx) {
  // c <= 3; constant indices keep idx in registers
#pragma unroll
  for (int t = 3; t >= 1; --t) {
    if ((u32)t > c) continue;
    const u64 x = colex_top(r, (u32)t, n);
    idx[t - 1] = (u32)x;
    r -= binom(x, t);
    n = x;
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
  // Guard table (guard_plan / guard_scan below), filled after construction:
  u32 gmode;          // GUARD_NONE, GUARD_TABLE or GUARD_ALLMAX
  u32 glog;           // log2 of the table's slots
  u32 gmin;           // lower bound of m(w) over every w of side A outside L
  u32 s1cap;          // guard lists longer than this take the plain lists instead
  u64 gtab;           // address of the table
};

// Guard tables.  For a family with free list Fr and nch chosen vertices per
// rank (fixed by its count vector), let m(w) = |Fr \ N(w)| for w in side A
// outside L.  A w with m(w) > k can never be added.  The guard table is an
// open-addressing set (load <= 1/2, linear probing from nh_slot) of the words
// w | m(w) << 30 for every w adjacent to a free vertex with m(w) <= k:
//  - nfree >= k + 1: every w with m(w) <= k neighbours one of the k + 1
//    lightest free vertices (fpick); a w absent from the table has m(w) > k;
//  - nfree <= k: the free vertices' lists give every w adjacent to one of them;
//    a w absent from the table ("outsider") has m(w) = nfree.
// With the table, a rank's candidates are the neighbours of its chosen
// vertices (guard_segments): an addable w with no chosen neighbour would miss
// m(w) + nch <= k members and no saturated one, for every rank of the family,
// and such a family is dropped whole (dominated, total 0) when its table is
// built.  This mirrors the reference's special-case check, which counts each
// candidate's intersection with conY (the free list) once per tuple and then
// tests only varY (the chosen vertices) per combination
// (include/for_checkmaximality.cuh:2587-2650).
enum : u32 { GUARD_NONE = 0, GUARD_TABLE = 1, GUARD_ALLMAX = 2 };
// A family's table build is a short serial chain per lane (it adds to the
// latency of the launch that built the family): it scans at most
// GUARD_CAND_MAX list entries and makes at most GUARD_WORK_MAX adjacency tests
// per lane, and is skipped for families of fewer than GUARD_MIN_RANKS
// combinations, which gain little from it.
constexpr u32 GUARD_CAND_MAX = 1024;
constexpr u32 GUARD_WORK_MAX = 64;
constexpr u64 GUARD_MIN_RANKS = 1024;
constexpr u32 GUARD_ABSENT = 0xFFFFFFFFu;
constexpr u32 GUARD_EMPTY = 0xFFFFFFFFu;
constexpr u32 GUARD_VMASK = (1u << 30) - 1u;  // vertex part of a table word (vertex IDs < 2^30)

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
// (its hash, when w has one).  miss0 misses are already counted; with
// `free_known` they are w's misses on the whole free list (from the family's
// guard table) and only the chosen vertices are looked up.
// NMAX bounds nch at compile time so the chosen vertices stay in registers.
template <int NMAX>
GMK_HD bool addable(const Graph& g, u32 w, const u32* freel, u32 nfree, const u32* ch, const u32* ctsz, u32 nch,
                    u32 k, u32 known, u32 miss0, bool free_known, Cost& cost) {
  const u32 ow = g.off[w], dw = g.off[w + 1] - ow;
  const u32* nw = g.adj + ow;
  cost.bytes += 8;
  // w needs at least (members still to look up) - (k - miss0) neighbours
  const u32 rs = (free_known ? 0u : nfree) + nch;
  if ((u64)dw + k < (u64)rs + miss0) return false;
  const u32* hw = nh_table(g, w, dw, cost);
  u32 miss = miss0;
  bool out = false;  // a saturated member is missed, or more than k members
  // a run-time loop (members fetched by a select over constant indices) keeps
  // one copy of the lookup code
  for (u32 i = 0; i < nch && !out; ++i) {
    u32 x = 0, ts = 0;
#pragma unroll
    for (int q = 0; q < NMAX; ++q)
      if ((u32)q == i) { x = ch[q]; ts = ctsz[q]; }
    if (x != known && !adjacent(hw, nw, dw, x, cost)) out = ts >= k || ++miss > k;
  }
  if (out) return false;
  if (free_known) return true;
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

// Misses of w on the free list, counted up to cap + 1; `tests` counts the
// adjacency tests made.
GMK_HD u32 free_misses(const Graph& g, u32 w, const u32* freel, u32 nfree, u32 cap, u32& tests, Cost& cost) {
  const u32 ow = g.off[w], dw = g.off[w + 1] - ow;
  cost.bytes += 8;
  if ((u64)dw + cap < nfree) return cap + 1;  // fewer than nfree - cap neighbours
  const u32* hw = nh_table(g, w, dw, cost);
  u32 m = 0, i = 0;
  for (; i < nfree && m <= cap; ++i) {
    ++tests;
    if (!adjacent(hw, g.adj + ow, dw, freel[i], cost)) ++m;
  }
  cost.bytes += 4ull * i;
  return m;
}

// m(w) from a family's guard table (size 2^lg), or GUARD_ABSENT.
GMK_HD u32 guard_find(const u32* t, u32 lg, u32 w, Cost& cost) {
  if (!t) return GUARD_ABSENT;
  const u32 mask = (1u << lg) - 1u;
  for (u32 s = nh_slot(w, lg);; s = (s + 1) & mask) {
    const u32 e = t[s];
    cost.bytes += 4;
    if (e == GUARD_EMPTY) return GUARD_ABSENT;
    if ((e & GUARD_VMASK) == w) return e >> 30;
  }
}

// What a family's guard build scans: the lists of np vertices (the k + 1
// lightest free ones, or every free one when nfree <= k), ctot entries in all.
struct GuardPlan {
  u32 poff[4];  // list of each scanned vertex: CSR offset and length
  u32 plen[4];
  u32 np;
  u32 ctot;
  u32 nch;      // chosen vertices per rank
  u32 sat;      // 1 when some chosen type is saturated (misses k members)
  u32 ok;       // 1 when the build may run (ctot <= GUARD_CAND_MAX, IDs fit the words)
};

GMK_HD void guard_plan(const Graph& g, const Family& f, u32 k, u64 min_ranks, GuardPlan& p, Cost& cost) {
  p.nch = 0;
  p.sat = 0;
  p.np = 0;
  p.ctot = 0;
  p.ok = 0;
  if (f.total < min_ranks) return;
  for (u32 t = 0; t < f.ntypes; ++t) {
    p.nch += f.cnt[t];
    if ((u32)popc(f.mask[t]) >= k) p.sat = 1;
  }
  const u32 np = f.nfree >= k + 1 ? k + 1 : f.nfree;  // k + 1 <= 4, nfree <= k <= 3
  u64 tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    p.poff[i] = 0;
    p.plen[i] = 0;
    if ((u32)i < np) {
      const u32 y = f.nfree >= k + 1 ? f.fpick[i] : f.freep[i];
      p.poff[i] = g.off[y];
      p.plen[i] = g.off[y + 1] - p.poff[i];
      tot += p.plen[i];
    }
  }
  cost.bytes += 12ull * np;
  p.np = np;
  p.ok = tot <= GUARD_CAND_MAX && (u64)g.a_hi <= GUARD_VMASK && (u64)g.b_hi <= GUARD_VMASK;
  p.ctot = p.ok ? (u32)tot : 0;
}

// Slots of a family's table: 2^lg >= 2 ctot (at least 64), so the load stays
// at most 1/2.
GMK_HD u32 guard_log2(u32 ctot) {
  u32 lg = 6;
  while ((1ull << lg) < 2ull * ctot) ++lg;
  return lg;
}

// One worker's share (entries cand0, cand0 + cstep, ...) of a family's guard
// build: m(w) of every candidate w outside L, passed to ins(w, m) when m <= k
// (always, when nfree <= k).  Returns the least m passed (k + 1: none); sets
// `over` when the worker's adjacency tests pass GUARD_WORK_MAX (the build
// then gives up and the family keeps the plain check).
template <class Ins>
GMK_HD u32 guard_scan(const Graph& g, const Family& f, u32 k, const GuardPlan& p, u32 cand0, u32 cstep, Ins& ins,
                      bool& over, Cost& cost) {
  u32 best = k + 1, tests = 0;
  for (u32 e = cand0; e < p.ctot; e += cstep) {
    // list and position of entry e (constant indices keep the plan in registers)
    u32 o = e, at = 0;
    bool found = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (!found && (u32)q < p.np) {
        if (o < p.plen[q]) { at = p.poff[q] + o; found = true; }
        else o -= p.plen[q];
      }
    }
    const u32 w = g.adj[at];
    cost.bytes += 4;
    if (in_tuple(f.L, f.j, w)) continue;
    const u32 m = free_misses(g, w, f.freep, f.nfree, k, tests, cost);
    if (tests > GUARD_WORK_MAX) { over = true; break; }
    if (m <= k) {
      ins(w, m);
      if (m < best) best = m;
    }
  }
  return best;
}

// Inserts w | m << 30 into a table of 2^lg slots (sequential form; the device
// build claims slots by compare-and-swap).
GMK_HD void guard_insert_seq(u32* t, u32 lg, u32 w, u32 m) {
  const u32 mask = (1u << lg) - 1u;
  for (u32 s = nh_slot(w, lg);; s = (s + 1) & mask) {
    if (t[s] == GUARD_EMPTY) { t[s] = w | (m << 30); return; }
    if ((t[s] & GUARD_VMASK) == w) return;
  }
}

// The family's guard mode from its scan (gmin_t: least m in the table, k + 1
// for none).  Sets `dominated` for a family with a w of m(w) + nch <= k and no
// saturated type (w can be added to every R of the family: the family is
// dropped), and `gmin` to a lower bound of m(w) over side A outside L.
GMK_HD u32 guard_finish(const Graph& g, u32 j, u32 nfree, u32 k, const GuardPlan& p, u32 gmin_t, bool& dominated,
                        u32& gmin) {
  dominated = false;
  gmin = 0;
  if (!p.sat) {
    if (gmin_t + p.nch <= k) { dominated = true; return GUARD_NONE; }
    if (nfree <= k && nfree + p.nch <= k) {
      // outsiders have m = nfree: they dominate, and they exist when side A
      // has more vertices than L and the scanned lists together
      dominated = (u64)(g.a_hi - g.a_lo) > (u64)j + p.ctot;
      return GUARD_NONE;
    }
  }
  if (nfree >= k + 1 && gmin_t > k) return GUARD_ALLMAX;  // no w can miss at most k members
  const u32 floor_m = nfree >= k + 1 ? k + 1 : nfree;
  gmin = gmin_t < floor_m ? gmin_t : floor_m;
  return GUARD_TABLE;
}

// The plain lists of a rank are at most as long as the k + 1 lightest free
// vertices' lists together (the scanned lists when nfree >= k + 1); with
// nfree <= k they include chosen vertices' lists anyway, and the guard lists
// are always taken.
GMK_HD u32 guard_s1cap(u32 nfree, u32 k, const GuardPlan& p) { return nfree >= k + 1 ? p.ctot : 0xFFFFFFFFu; }

// Sequential guard build of one family (host model): table from the arena.
template <class Arena>
GMK_HD void guard_build(const Graph& g, Family& f, u32 k, Arena& ar, Cost& cost, u64 min_ranks = GUARD_MIN_RANKS) {
  f.gmode = GUARD_NONE;
  f.gtab = 0;
  f.glog = 0;
  f.gmin = 0;
  if (f.total == 0) return;
  GuardPlan p;
  guard_plan(g, f, k, min_ranks, p, cost);
  if (!p.ok) return;
  const u32 lg = guard_log2(p.ctot);
  u32* t = nullptr;
  if (p.ctot) {
    const u64 at = ar.alloc(1ull << lg);
    if (at == SAT) return;
    t = ar.ptr(at);
    for (u64 s = 0; s < (1ull << lg); ++s) t[s] = GUARD_EMPTY;
  }
  bool over = false;
  auto ins = [&](u32 w, u32 m) { guard_insert_seq(t, lg, w, m); };
  const u32 gm = guard_scan(g, f, k, p, 0, 1, ins, over, cost);
  if (over) return;
  bool dominated = false;
  const u32 mode = guard_finish(g, f.j, f.nfree, k, p, gm, dominated, f.gmin);
  if (dominated) f.total = 0;
  f.gmode = mode;
  f.s1cap = guard_s1cap(f.nfree, k, p);
  if (mode == GUARD_TABLE) { f.gtab = (u64)t; f.glog = lg; }
}

// What the checks read of a family: pointers into its record (the tuple, the
// free list, the lightest free vertices) and scalars, so the record's arrays are
// read from memory where they are used instead of being held in registers.
struct FamRef {
  const u32* L;
  u32 j;
  const u32* freep;
  u32 nfree;
  const u32* fpick;
  u32 nfpick;
  const u32* gtab;
  u32 glog, gmode, gmin, s1cap;
};
GMK_HD FamRef fam_ref(const Family& f) {
  return FamRef{f.L, f.j, f.freep, f.nfree, f.fpick, f.nfpick, (const u32*)f.gtab, f.glog, f.gmode, f.gmin, f.s1cap};
}

// Candidates for a vertex that could join (L, R): up to 4 lists of vertices of
// side A (neighbour lists), each with a member already known to be adjacent to
// its entries, and optionally a vertex every candidate must be adjacent to.
// `decided` >= 0 when no candidate is needed (1 = maximal).  `guard`: the
// lists are chosen vertices' lists and candidates are tested with the
// family's guard table (guard_joins).
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
  u32 guard;
};

// Candidate lists of a rank from its chosen vertices and the family's guard
// table.  Returns false (cs untouched) when they hold more entries than
// f.s1cap, the length of the plain lists' free part (the rank then takes the
// plain lists).
//  - With h chosen neighbours an addable w misses m(w) + nch - h <= k members,
//    so h >= hmin = max(1, nch - k + gmin, saturated chosen), and w neighbours
//    one of any nch - hmin + 1 chosen vertices (<= k + 1 - gmin <= 4 lists):
//    the lightest ones are taken;
//  - with a saturated chosen vertex, its list alone also holds every
//    candidate (the shorter choice is taken).
template <int NMAX>
GMK_HD bool guard_segments(const Graph& g, const FamRef& f, const u32* ch, const u32* ctsz, u32 nch, u32 k,
                           CandSegs& cs, Cost& cost) {
  u32 nsat = 0;
#pragma unroll
  for (int i = 0; i < NMAX; ++i) nsat += ((u32)i < nch && ctsz[i] >= k) ? 1u : 0u;
  int hmin = (int)nch - (int)k + (int)f.gmin;
  if (hmin < 1) hmin = 1;
  if (hmin < (int)nsat) hmin = (int)nsat;
  if (hmin > (int)nch) {  // no w has enough chosen neighbours
    cs.n = 0;
    cs.decided = 1;
    return true;
  }
  const u32 q = nch - (u32)hmin + 1;  // <= 4
  // the q chosen vertices of least degree (4 slots by constant index), and the
  // saturated one of least degree
  u32 pk[4] = {0, 0, 0, 0}, pd[4] = {0, 0, 0, 0}, po[4] = {0, 0, 0, 0};
  u32 np = 0, si = 0xFFFFFFFFu, sd = 0xFFFFFFFFu, so = 0;
#pragma unroll
  for (int i = 0; i < NMAX; ++i) {
    if ((u32)i < nch) {
      const u32 y = ch[i], oy = g.off[y], d = g.off[y + 1] - oy;
      cost.bytes += 8;
      if (ctsz[i] >= k && d < sd) { sd = d; si = y; so = oy; }
      if (np < q) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if ((u32)r == np) { pk[r] = y; pd[r] = d; po[r] = oy; }
        ++np;
      } else {
        u32 worst = 0, wd = pd[0];
#pragma unroll
        for (int r = 1; r < 4; ++r)
          if ((u32)r < np && pd[r] > wd) { worst = r; wd = pd[r]; }
        if (d < wd) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if ((u32)r == worst) { pk[r] = y; pd[r] = d; po[r] = oy; }
        }
      }
    }
  }
  u64 tot = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) tot += (u32)r < np ? pd[r] : 0u;
  if (si != 0xFFFFFFFFu && sd < tot) {
    np = 1;
    pk[0] = si;
    pd[0] = sd;
    po[0] = so;
    tot = sd;
  }
  if (tot > f.s1cap) return false;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const bool on = (u32)r < np;
    cs.lst[r] = g.adj + (on ? po[r] : 0u);
    cs.len[r] = on ? pd[r] : 0u;
    cs.known[r] = on ? pk[r] : 0xFFFFFFFFu;
  }
  cs.n = np;
  cs.guard = 1;
  return true;
}

// The candidate lists of (L, R) for family f and chosen vertices ch (nch, with
// their type sizes ctsz).
//  - Guard lists (guard_segments), when the family has a guard table and they
//    are short enough.
//  - Plain lists.  With a saturated member: an addable w is adjacent to the
//    saturated member `best` of least degree and to all but k members of R, so
//    it neighbours one of any k + 1 other members.  When the k + 1 other
//    members of least degree have fewer neighbours in all than `best`, their
//    lists give the candidates (kept if adjacent to `best`); otherwise N(best)
//    does.  With none: an addable w misses at most k members of R, so it is
//    adjacent to one of any k + 1 members; the k + 1 members of least degree
//    among the chosen vertices and fpick give the candidates.
template <int NMAX>
GMK_HD void candidate_segments(const Graph& g, const FamRef& f, const u32* ch, const u32* ctsz, u32 nch, u32 k,
                               CandSegs& cs, Cost& cost) {
  cs.n = 0;
  cs.must = 0xFFFFFFFFu;
  cs.must_nb = nullptr;
  cs.must_nh = nullptr;
  cs.must_deg = 0;
  cs.decided = -1;
  cs.guard = 0;
  if (f.gmode == GUARD_ALLMAX) {  // no vertex of side A misses at most k free vertices
    cs.decided = 1;
    return;
  }
  if (f.gmode == GUARD_TABLE && guard_segments<NMAX>(g, f, ch, ctsz, nch, k, cs, cost)) return;
  const u32 nfree = f.nfree, rs = nfree + nch;
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
    cs.decided = (u64)(g.a_hi - g.a_lo) <= f.j ? 1 : 0;
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
  for (u32 i = 0; i < f.nfpick; ++i) take(f.fpick[i]);
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
// With guard lists, w's misses on the free list come from the family's guard
// table (absent: nfree when nfree <= k, else more than k) and only the chosen
// vertices are looked up.
template <int NMAX>
GMK_HD bool candidate_joins(const Graph& g, const FamRef& f, const u32* ch, const u32* ctsz, u32 nch, u32 k,
                            const CandSegs& cs, u32 known, u32 w, Cost& cost) {
  if (in_tuple(f.L, f.j, w)) return false;
  u32 miss0 = 0;
  if (cs.guard) {
    miss0 = guard_find(f.gtab, f.glog, w, cost);
    if (miss0 == GUARD_ABSENT) {
      if (f.nfree > k) return false;
      miss0 = f.nfree;
    }
  } else if (cs.must != 0xFFFFFFFFu && !adjacent(cs.must_nh, cs.must_nb, cs.must_deg, w, cost)) {
    return false;
  }
  return addable<NMAX>(g, w, f.freep, f.nfree, ch, ctsz, nch, k, known, miss0, cs.guard != 0, cost);
}

// 1 if (L, R) is maximal on side A, 0 if not; -1 when the first `budget`
// candidates (in list order) cannot join and more remain: the caller tests the
// rest another way.
template <int NMAX>
GMK_HD int side_a_maximal(const Graph& g, const FamRef& f, const u32* ch, const u32* ctsz, u32 nch, u32 k, Cost& cost,
                          u32 budget = 0xFFFFFFFFu) {
  CandSegs cs;
  candidate_segments<NMAX>(g, f, ch, ctsz, nch, k, cs, cost);
  if (cs.decided >= 0) return cs.decided;
  u32 tested = 0;
  // at most 4 lists, fetched by a select over constant indices (cs stays in
  // registers, and the candidate test is emitted once)
  for (u32 q = 0; q < cs.n; ++q) {
    const u32* lst = cs.lst[0];
    u32 len = 0, known = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if ((u32)s == q) { lst = cs.lst[s]; len = cs.len[s]; known = cs.known[s]; }
    for (u32 e = 0; e < len; ++e) {
      if (tested++ == budget) return -1;
      cost.bytes += 4;
      if (candidate_joins<NMAX>(g, f, ch, ctsz, nch, k, cs, known, lst[e], cost)) return 0;
    }
  }
  return 1;
}

}  // namespace thin
}  // namespace gmk
