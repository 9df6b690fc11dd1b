// Branch-and-bound state expansion for maximal k-biplex enumeration on gfx950.
//
// One wave64 expands one search state.  A state lives in HBM as bitsets over the
// local vertex IDs of its partition:
//     [header][S_L C_L X_L][S_R C_R X_R]      (each set W words of 64 bits)
// S = vertices taken, C = candidates that can still be added, X = vertices that
// can be added to S but are excluded from this subtree.  Invariant: for every
// v in C or X, S + v is a k-biplex.  A state with C = X = {} and both sides of S
// at least lb is a maximal k-biplex of its partition.
//
// The branching and pruning rules are those of BPPivot's branchNew
// (baselines/MaximalBiPlex/src/biplex/biplex.cpp:781-1082): one-pass degree
// pruning of C (biplex.cpp:695-713), the size bound, the non-degree bound
// (biplex.cpp:840-866), the binary branch on a candidate missed by the S vertex
// with most misses (biplex.cpp:912-965) and otherwise pivot branching
// (biplex.cpp:970-1071).  The per-vertex counters BPPivot keeps incrementally
// (degC, nonNbrS) are recomputed here from the bitsets with AND + popcount over
// adjacency rows, which is what the wave does in parallel: lane i owns vertices
// i, i+64, ...  Ties in every argmax/argmin go to the smallest local ID.
//
// Widths W up to 128 words (8192 vertices per side) keep the wave's working
// set (sets, per-vertex counters) in LDS; wider partitions, up to 1024 words
// (65536 per side), keep it in an HBM slice of their own per wave
// (bb_expand_wide), with the same code.
//
// Children of a state are written to a contiguous block claimed with one
// atomicAdd; when the block would cross the buffer's end, the claimed slots
// inside the buffer are marked as holes (skipped when read) and the parent is
// written to the retry list unchanged, to be expanded again in a later batch.
#include "gmk_device.h"

#include <type_traits>

namespace gmk {

struct Stats {
  u64 states, intersections, bytes, leaves;
};

// Per-vertex degree counters: 16 bits up to 512 words (32768 vertices per side).
template <int W>
using deg_t = typename std::conditional<(W > 512), u32, uint16_t>::type;

// Widest partition whose working set lives in LDS.
constexpr int W_LDS_MAX = 128;

template <int W>
struct __align__(16) Smem {
  u64 set[6 * W];         // S_L C_L X_L S_R C_R X_R
  u64 full[2][W];         // members with ms >= k
  u64 k1[2][W];           // members with ms == k - 1
  u64 bad[2][W];          // S members with non-degree >= k (blocks X pivots)
  u64 cand[W];            // pivot candidates (on side cand_side)
  u64 sat[W];             // S members saturated by the include being built
  u64 andsat[W];          // AND of the rows of sat
  u64 prefix[W];           // pivot candidates already branched on
  deg_t<W> deg[2][64 * W];
  uint8_t ms[2][64 * W];
  uint16_t widx[W];        // words of the opposite side that the current pass reads
  u64 rm[2][W];            // candidates removed by the degree pruning of this state
  u32 c1[64 * W];          // child bound: neighbours of the include missing only this S member
};

// Lists in sm.widx the words w < wb with mask_a[w] | mask_b[w] != 0; returns
// their number.  Words outside the list contribute nothing to the counters.
template <int W>
__device__ __forceinline__ u32 word_list(Smem<W>& sm, const u64* ma, const u64* mb, u32 wb) {
  const int lane = threadIdx.x & 63;
  u32 n = 0;
  for (u32 base = 0; base < wb; base += 64) {
    const u32 w = base + lane;
    const bool nz = w < wb && ((ma[w] | (mb ? mb[w] : 0)) != 0);
    const u64 bal = __ballot(nz);
    if (nz) sm.widx[n + __popcll(bal & ((1ull << lane) - 1ull))] = (uint16_t)w;
    n += (u32)__popcll(bal);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  return n;
}

// End of a counter pass: LDS working sets need only the wave barrier; the
// HBM ones of wide partitions also wait for the wave's stores.
template <int W>
__device__ __forceinline__ void pass_barrier() {
  if constexpr (W > W_LDS_MAX) wave_barrier();
  else __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ u64* setp(u64* base, int side, int which, int W) {
  return base + (side * 3 + which) * W;
}

// Stats for all members of side a: deg = |N(v) & C_b|, ms = |S_b \ N(v)|.
template <int W>
__device__ void side_stats(Smem<W>& sm, const PartDesc& P, const u64* rows, int a, Stats& st) {
  const int lane = threadIdx.x;
  const u32 na = a ? P.nr : P.nl;
  const u32 wa = a ? P.wr : P.wl;
  const u32 wb = a ? P.wl : P.wr;
  const u64* rowsA = rows + (a ? P.rows_r : P.rows_l);
  const u64* Sa = setp(sm.set, a, 0, W);
  const u64* Ca = setp(sm.set, a, 1, W);
  const u64* Xa = setp(sm.set, a, 2, W);
  const u64* Sb = setp(sm.set, a ^ 1, 0, W);
  const u64* Cb = setp(sm.set, a ^ 1, 1, W);
  const u32 nw = word_list<W>(sm, Cb, Sb, wb);
  u32 tests = 0;
  for (u32 i = 0; i < wa; ++i) {
    const u64 memb = Sa[i] | Ca[i] | Xa[i];
    if (memb == 0) continue;  // the counters of non-members are never read
    const u32 v = i * 64 + lane;
    const bool m = v < na && ((memb >> lane) & 1ull);
    u32 d = 0, s = 0;
    if (m) {
      const u64* row = rowsA + (u64)v * wb;
      for (u32 q = 0; q < nw; ++q) {
        const u32 w = sm.widx[q];
        const u64 r = row[w];
        d += __popcll(r & Cb[w]);
        s += __popcll(Sb[w] & ~r);
      }
      ++tests;
    }
    sm.deg[a][v] = (deg_t<W>)d;
    sm.ms[a][v] = (uint8_t)(s > 255 ? 255 : s);
  }
  st.intersections += tests;
  st.bytes += (u64)tests * nw * 8;
  pass_barrier<W>();  // widx is rewritten by the next pass
}

// deg of the members of side a after the vertices rm_b left C_b:
// deg -= |N(v) & rm_b|, read over the (few) words rm_b occupies.
template <int W>
__device__ void side_deg_sub(Smem<W>& sm, const PartDesc& P, const u64* rows, int a, Stats& st) {
  const int lane = threadIdx.x;
  const u32 na = a ? P.nr : P.nl;
  const u32 wa = a ? P.wr : P.wl;
  const u32 wb = a ? P.wl : P.wr;
  const u64* rowsA = rows + (a ? P.rows_r : P.rows_l);
  const u64* Sa = setp(sm.set, a, 0, W);
  const u64* Ca = setp(sm.set, a, 1, W);
  const u64* Xa = setp(sm.set, a, 2, W);
  const u64* Rb = sm.rm[a ^ 1];
  const u32 nw = word_list<W>(sm, Rb, nullptr, wb);
  if (nw == 0) return;
  u32 tests = 0;
  for (u32 i = 0; i < wa; ++i) {
    const u64 memb = Sa[i] | Ca[i] | Xa[i];
    if (memb == 0) continue;
    const u32 v = i * 64 + lane;
    if (v < na && ((memb >> lane) & 1ull)) {
      const u64* row = rowsA + (u64)v * wb;
      u32 d = 0;
      for (u32 q = 0; q < nw; ++q) {
        const u32 w = sm.widx[q];
        d += __popcll(row[w] & Rb[w]);
      }
      sm.deg[a][v] = (deg_t<W>)(sm.deg[a][v] - d);
      ++tests;
    }
  }
  st.intersections += tests;
  st.bytes += (u64)tests * nw * 8;
  pass_barrier<W>();
}

template <int W>
__device__ __forceinline__ u32 popc_set(const u64* s, u32 w) {
  u32 c = 0;
  for (u32 i = 0; i < w; ++i) c += __popcll(s[i]);
  return c;
}

// Emit S as a result record (global IDs ascending) and fold its fingerprint.
template <int W>
__device__ void emit_state(Smem<W>& sm, const PartDesc& P, const u32* maps, const ExpandArgs& a) {
  const int lane = threadIdx.x;
  const u64* SL = setp(sm.set, 0, 0, W);
  const u64* SR = setp(sm.set, 1, 0, W);
  const u32 nl = popc_set<W>(SL, P.wl), nr = popc_set<W>(SR, P.wr);
  // fingerprint first: it decides whether the record is a duplicate
  u64 h = 0, h2 = 0;
  for (int side = 0; side < 2; ++side) {
    const u64* Sx = side ? SR : SL;
    const u32 wx = side ? P.wr : P.wl;
    const u32* map = maps + (side ? P.map_r : P.map_l);
    for (u32 i = 0; i < wx; ++i)
      if ((Sx[i] >> lane) & 1ull) {
        const u32 id = map[i * 64 + lane];
        h += mix64(id);
        if (a.dedup.on) h2 += mix64b(id);  // the dedup table's check word
      }
  }
  h = wave_sum_u64(h);
  const u64 fp = fmix64(h + (u64)nl * FP_A + (u64)nr * FP_B);
  if (a.dedup.on) h2 = fmix64(wave_sum_u64(h2) + (u64)nl * FP_B + (u64)nr * FP_A);
  u64 slot = 0;
  if (lane == 0) {
    bool take = dedup_insert(a.dedup, fp, h2, a.flags, blockIdx.x);
    if (take && a.gcount) {
      // per-graph cap and digest
      const unsigned long long gs = atomicAdd(&a.gcount[P.graph], 1ull);
      if (gs >= a.gcap) take = false;
      else {
        atomicAdd((unsigned long long*)&a.gfp[2 * P.graph], (unsigned long long)fp);
        atomicXor((unsigned long long*)&a.gfp[2 * P.graph + 1], (unsigned long long)fp);
      }
    }
    slot = take ? atomicAdd((unsigned long long*)a.res_count, 1ull) : ~0ull;
  }
  slot = wave_bcast_u64(slot, 0);
  if (slot == ~0ull) return;  // duplicate, or the graph's cap is reached
  if (slot >= a.limit) {
    if (lane == 0) atomicOr(a.flags, FLAG_LIMIT);
    return;
  }
  if (slot + 1 == a.limit && lane == 0) atomicOr(a.flags, FLAG_LIMIT);
  const u64 len = 2ull + nl + nr;
  u64 pos = 0;
  if (a.keep && lane == 0) pos = atomicAdd((unsigned long long*)a.res_words, (unsigned long long)len);
  pos = wave_bcast_u64(pos, 0);
  const bool write = a.keep && pos + len <= a.res_cap;
  if (a.keep && !write && lane == 0) atomicOr(a.flags, FLAG_RES_OVERFLOW);
  if (write) {
    for (int side = 0; side < 2; ++side) {
      const u64* Sx = side ? SR : SL;
      const u32 wx = side ? P.wr : P.wl;
      const u32* map = maps + (side ? P.map_r : P.map_l);
      const u64 off = pos + 1 + (side ? 1 + nl : 0);
      u32 before = 0;
      for (u32 i = 0; i < wx; ++i) {
        const u64 word = Sx[i];
        if ((word >> lane) & 1ull) {
          const u32 rank = before + __popcll(word & ((1ull << lane) - 1ull));
          a.res[off + rank] = map[i * 64 + lane];
        }
        before += __popcll(word);
      }
    }
  }
  if (lane == 0) {
    if (write) {
      a.res[pos] = nl;
      a.res[pos + 1 + nl] = nr;
    }
    u64* fs = a.fp + (blockIdx.x % NSTRIPE) * 2;
    atomicAdd((unsigned long long*)&fs[0], (unsigned long long)fp);
    atomicXor((unsigned long long*)&fs[1], (unsigned long long)fp);
  }
}

// Writes one child: the parent with `moved` (on side mv_side) moved from C to
// X, then (when inc != NONE) vertex inc of side us included.  A child that
// fails its bounds (below) is written as a hole (partition INVALID_PID), which
// the next expansion skips; returns whether it passed.
template <int W>
__device__ bool write_child(Smem<W>& sm, const PartDesc& P, const u64* rows, u32 k, u64* out, u32 pid,
                            u32 inc, int us, int mv_side, const u64* moved /* W words or null */,
                            u32 moved_single /* used when moved == null, NONE = none */) {
  const int lane = threadIdx.x;
  constexpr int NWS = 6 * W;
  const int os = us ^ 1;
  const u32 wos = os ? P.wr : P.wl;
  const u64* rowU = nullptr;
  bool ufull = false;
  if (inc != NONE) {
    rowU = rows + (us ? P.rows_r : P.rows_l) + (u64)inc * (us ? P.wl : P.wr);
    ufull = sm.ms[us][inc] >= k;
    // saturated S members of the other side: ms == k-1 and not adjacent to inc
    for (u32 w = lane; w < (u32)W; w += 64) {
      u64 s = 0;
      if (w < wos && k > 0) s = setp(sm.set, os, 0, W)[w] & ~rowU[w] & sm.k1[os][w];
      sm.sat[w] = s;
    }
    wave_barrier();
    // AND of the saturated members' rows (over side us)
    const u32 wus = us ? P.wr : P.wl;
    for (u32 w = lane; w < (u32)W; w += 64) {
      u64 acc = ~0ull;
      for (u32 i = 0; i < wos; ++i) {
        u64 s = sm.sat[i];
        while (s) {
          const u32 b = __ffsll((long long)s) - 1;
          s &= s - 1;
          const u32 v = i * 64 + b;
          if (w < wus) acc &= rows[(os ? P.rows_r : P.rows_l) + (u64)v * wus + w];
        }
      }
      sm.andsat[w] = acc;
    }
    wave_barrier();
  }
  u32 sc0 = 0, sc1 = 0;  // |S| + |C| of the child per side
  for (int j = lane; j < NWS; j += 64) {
    const int side = j / (3 * W);
    const int which = (j / W) % 3;
    const int w = j % W;
    u64 val = sm.set[j];
    // moves C -> X
    if (side == mv_side && which != 0) {
      u64 mv = 0;
      if (moved) mv = moved[w];
      else if (moved_single != NONE && (moved_single >> 6) == (u32)w) mv = 1ull << (moved_single & 63);
      val = which == 1 ? (val & ~mv) : (val | mv);
    }
    if (inc != NONE) {
      if (side == os && which != 0) {
        const u64 r = (u32)w < wos ? rowU[w] : 0;
        const u64 kill = ~r & (ufull ? ~0ull : sm.full[os][w]);
        val &= ~kill;
      }
      if (side == us) {
        const u64 bit = ((inc >> 6) == (u32)w) ? (1ull << (inc & 63)) : 0;
        if (which == 0) val |= bit;
        else val &= ~bit & sm.andsat[w];
      }
    }
    if (which != 2) {
      if (side == 0) sc0 += __popcll(val);
      else sc1 += __popcll(val);
    }
    if (out) out[1 + j] = val;
  }
  // Bounds of the child (biplex.cpp:949-951, 1040-1042, 1054-1056): S + C must
  // reach lb on both sides, and with an include u the opposite side must be
  // able to reach lb.  That side gains at most k - ms(u) non-neighbours of u,
  // and neighbours w of u in C each missing ms(w) members of S + u; those
  // misses draw on the members' remaining budgets k - ms.  Counted as
  // biplexUpperBound (biplex.cpp:556-603) counts them: every w missing no
  // member, per member x at most its budget of the w missing only x, then w
  // missing two or more while their misses fit what is left of the budgets'
  // sum (where biplexUpperBound also checks one member per w, so this never
  // bounds below it).
  const u32 lb = P.lb;
  if (lb <= 1) {  // every side of S + C is non-empty in any state that can emit: nothing to bound
    if (out && lane == 0) out[0] = pid;
    return true;
  }
  bool keep = (u32)wave_sum_u64(sc0) >= lb && (u32)wave_sum_u64(sc1) >= lb;
  const u32 ms_u = inc != NONE ? sm.ms[us][inc] : 0;
  const u32 nso = inc != NONE ? popc_set<W>(setp(sm.set, os, 0, W), wos) : 0;
  // the bound only matters when |S| and u's own budget do not reach lb already
  if (keep && inc != NONE && nso + (k > ms_u ? k - ms_u : 0) < lb) {
    const u64* Sus = setp(sm.set, us, 0, W);
    const u64* Cos = setp(sm.set, os, 1, W);
    const u32 wus = us ? P.wr : P.wl, nus = us ? P.nr : P.nl, nos = os ? P.nr : P.nl;
    u64 ub = (k > ms_u ? k - ms_u : 0);
    u64 sup = ub;  // u's own budget
    for (u32 i = 0; i < wus; ++i) {
      const u32 v = i * 64 + lane;
      if (v < nus && ((Sus[i] >> lane) & 1ull)) sup += k - sm.ms[us][v];
    }
    sup = wave_sum_u64(sup);
    ub += nso;
    // neighbours of u in C of the other side, counted by their ms (cost)
    const u64* mvos = (mv_side == os && moved) ? moved : nullptr;
    for (u32 cst = 0; cst <= k && ub < lb; ++cst) {
      if (cst == 1) {
        // each w missing exactly one member x of S uses one unit of x's budget
        // k - ms(x) (as biplexUpperBound, biplex.cpp:583-597): at most
        // min(count of such w, budget) of them per x
        for (u32 i = 0; i < (u32)W; ++i) sm.c1[i * 64 + lane] = 0;
        wave_barrier();
        const u64* rowsO = rows + (os ? P.rows_r : P.rows_l);
        for (u32 i = 0; i < wos; ++i) {
          const u64 nb = rowU[i] & Cos[i] & (mvos ? ~mvos[i] : ~0ull);
          const u32 v = i * 64 + lane;
          if (((nb >> lane) & 1ull) && v < nos && sm.ms[os][v] == 1) {
            const u64* rw = rowsO + (u64)v * wus;
            for (u32 q = 0; q < wus; ++q) {
              const u64 miss = Sus[q] & ~rw[q];
              if (miss) { atomicAdd(&sm.c1[q * 64 + (__ffsll((long long)miss) - 1)], 1u); break; }
            }
          }
        }
        wave_barrier();
        u64 take = 0;
        for (u32 i = 0; i < wus; ++i) {
          const u32 x = i * 64 + lane;
          if (x < nus && ((Sus[i] >> lane) & 1ull)) {
            const u32 bx = k - sm.ms[us][x];
            take += sm.c1[x] < bx ? sm.c1[x] : bx;
          }
        }
        take = wave_sum_u64(take);
        ub += take;
        sup -= take;  // take <= sum of the members' budgets <= sup
        wave_barrier();
        continue;
      }
      u32 cnt = 0;
      for (u32 i = 0; i < wos; ++i) {
        const u64 nb = rowU[i] & Cos[i] & (mvos ? ~mvos[i] : ~0ull);
        const u32 v = i * 64 + lane;
        cnt += __popcll(__ballot(((nb >> lane) & 1ull) && v < nos && sm.ms[os][v] == cst));
      }
      const u64 take = cst == 0 ? cnt : (sup / cst < cnt ? sup / cst : cnt);
      ub += take;
      sup -= take * cst;
      if (cst > 0 && sup < cst) break;
    }
    if (ub < lb) keep = false;
  }
  if (out && lane == 0) out[0] = keep ? pid : INVALID_PID;
  return keep;
}

// Expands state sidx of the batch; counters accumulate into st.
template <int W>
__device__ void expand_one(Smem<W>& sm, const ExpandArgs& a, u32 sidx, Stats& st) {
  const int lane = threadIdx.x & 63;
  constexpr int STRIDE = 1 + 6 * W;
  const u64* in = a.in + ((a.ctl ? a.ctl->in_off : 0ull) + sidx) * STRIDE;
  const u32 pid = (u32)in[0];
  if (pid == INVALID_PID) return;  // hole left by a child block that did not fit
  const PartDesc P = a.parts[pid];
  const u32 k = P.k;
  const u32 lb = P.lb;
  if (a.gcount && __hip_atomic_load(&a.gcount[P.graph], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= a.gcap)
    return;  // this graph already has all the results it keeps
  wave_barrier();  // the previous state of this wave is done with the LDS
  for (int j = lane; j < 6 * W; j += 64) sm.set[j] = in[1 + j];
  wave_barrier();
  st.states += 1;
  st.bytes += STRIDE * 8;

  side_stats<W>(sm, P, a.rows, 0, st);
  side_stats<W>(sm, P, a.rows, 1, st);
  wave_barrier();

  u32 nS[2], nC[2], nX[2];
  auto count_sets = [&]() {
    nS[0] = popc_set<W>(setp(sm.set, 0, 0, W), P.wl);
    nC[0] = popc_set<W>(setp(sm.set, 0, 1, W), P.wl);
    nX[0] = popc_set<W>(setp(sm.set, 0, 2, W), P.wl);
    nS[1] = popc_set<W>(setp(sm.set, 1, 0, W), P.wr);
    nC[1] = popc_set<W>(setp(sm.set, 1, 1, W), P.wr);
    nX[1] = popc_set<W>(setp(sm.set, 1, 2, W), P.wr);
  };
  count_sets();

  // One-pass degree pruning of C (biplex.cpp:695-713).  Decisions use the
  // degrees before any removal; removal only lowers degrees, so every removed
  // vertex is one the sequential pass would also be allowed to remove.
  bool pruned_any = false;
  for (int w = lane; w < W; w += 64) { sm.rm[0][w] = 0; sm.rm[1][w] = 0; }
  wave_barrier();
  for (int a_ = 0; a_ < 2; ++a_) {
    const u32 na = a_ ? P.nr : P.nl, wa = a_ ? P.wr : P.wl;
    u64* Ca = setp(sm.set, a_, 1, W);
    for (u32 i = 0; i < wa; ++i) {
      const u32 v = i * 64 + lane;
      bool inC = v < na && ((Ca[i] >> lane) & 1ull);
      bool rm = false;
      if (inC) {
        const int64_t d = (int64_t)sm.deg[a_][v] + ((int64_t)nS[a_ ^ 1] - (int64_t)sm.ms[a_][v]) + (int64_t)k;
        rm = d < (int64_t)lb;
      }
      const u64 rmw = __ballot(rm);
      wave_barrier();
      if (rmw) {
        if (lane == 0) { Ca[i] &= ~rmw; sm.rm[a_][i] = rmw; }
        pruned_any = true;
      }
      wave_barrier();
    }
  }
  if (pruned_any) {
    count_sets();
    side_deg_sub<W>(sm, P, a.rows, 0, st);
    side_deg_sub<W>(sm, P, a.rows, 1, st);
    wave_barrier();
  }

  u32 nchild = 0;
  bool emit = false;
  // kind: 0 = none, 1 = binary (include u / exclude u), 2 = pivot
  int kind = 0;
  u32 bin_u = NONE;
  int bin_side = 0;
  u32 pivot = NONE;
  int pside = 0;
  bool pivot_in_c = false;

  if (nS[0] + nC[0] < lb || nS[1] + nC[1] < lb) goto finish;
  if (nC[0] == 0 && nC[1] == 0) {
    if (nX[0] == 0 && nX[1] == 0 && nS[0] >= lb && nS[1] >= lb) emit = true;
    st.leaves = 1;
    goto finish;
  }
  {
    // Non-degree of S members: |C_b| - deg + ms (biplex.cpp:847-866).
    int64_t nd[2];
    u32 sp[2];
    for (int a_ = 0; a_ < 2; ++a_) {
      const u32 na = a_ ? P.nr : P.nl, wa = a_ ? P.wr : P.wl;
      const u64* Sa = setp(sm.set, a_, 0, W);
      u64 best = 0;  // packed (nd + 1) << 32 | ~id ; 0 = none
      for (u32 i = 0; i < wa; ++i) {
        const u32 v = i * 64 + lane;
        const bool m = v < na && ((Sa[i] >> lane) & 1ull);
        bool bad = false;
        if (m) {
          const u32 d = nC[a_ ^ 1] - sm.deg[a_][v] + sm.ms[a_][v];
          const u64 key = ((u64)(d + 1) << 32) | (u64)(0xFFFFFFFFu - v);
          best = key > best ? key : best;
          bad = d >= k;
        }
        const u64 bw = __ballot(bad);
        if (lane == 0) sm.bad[a_][i] = bw;
      }
      best = wave_max_u64(best);
      if (best == 0) { nd[a_] = INT64_MIN / 4; sp[a_] = NONE; }
      else { nd[a_] = (int64_t)(best >> 32) - 1; sp[a_] = 0xFFFFFFFFu - (u32)(best & 0xFFFFFFFFu); }
      if (nS[a_] > 0 && (int64_t)(nS[a_ ^ 1] + nC[a_ ^ 1]) - nd[a_] < (int64_t)lb - (int64_t)k) goto finish;
    }
    // full / k1 bitsets over all members (used by the include filter)
    for (int a_ = 0; a_ < 2; ++a_) {
      const u32 na = a_ ? P.nr : P.nl, wa = a_ ? P.wr : P.wl;
      for (u32 i = 0; i < (u32)W; ++i) {
        const u32 v = i * 64 + lane;
        const bool ok = i < wa && v < na;
        const u32 m = ok ? sm.ms[a_][v] : 0;
        const u64 f = __ballot(ok && m >= k);
        const u64 q = __ballot(ok && k > 0 && m == k - 1);
        if (lane == 0) { sm.full[a_][i] = f; sm.k1[a_][i] = q; }
      }
    }
    wave_barrier();

    if (nd[0] > (int64_t)k || nd[1] > (int64_t)k) {
      // Binary branch (biplex.cpp:912-965).
      const int side = nd[0] > nd[1] ? 0 : 1;
      const int os = side ^ 1;
      const u32 n_os = os ? P.nr : P.nl, w_os = os ? P.wr : P.wl;
      const u64* rowSp = a.rows + (side ? P.rows_r : P.rows_l) + (u64)sp[side] * w_os;
      const u64* Cos = setp(sm.set, os, 1, W);
      u64 best = ~0ull;  // minimise (d, 255 - ms, id)
      for (u32 i = 0; i < w_os; ++i) {
        const u32 v = i * 64 + lane;
        const u64 r = rowSp[i];
        const bool c = v < n_os && ((Cos[i] >> lane) & 1ull) && !((r >> lane) & 1ull);
        if (c) {
          const u32 d = sm.deg[os][v] + nS[side] - sm.ms[os][v];
          const u64 key = ((u64)d << 40) | ((u64)(255 - sm.ms[os][v]) << 32) | v;
          best = key < best ? key : best;
        }
      }
      best = wave_min_u64(best);
      if (best != ~0ull) {
        kind = 1;
        bin_u = (u32)(best & 0xFFFFFFFFu);
        bin_side = os;
        nchild = 2;
      }
    } else {
      // Pivot branching (biplex.cpp:970-1071).
      int64_t dp[2];
      u32 pv[2];
      bool pin[2];
      for (int a_ = 0; a_ < 2; ++a_) {
        const u32 na = a_ ? P.nr : P.nl, wa = a_ ? P.wr : P.wl;
        const u32 wb = a_ ? P.wl : P.wr;
        const u64* Ca = setp(sm.set, a_, 1, W);
        const u64* Xa = setp(sm.set, a_, 2, W);
        const u64* rowsA = a.rows + (a_ ? P.rows_r : P.rows_l);
        u64 bc = 0, bx = 0;  // (deg + 1) << 32 | ~id
        for (u32 i = 0; i < wa; ++i) {
          const u32 v = i * 64 + lane;
          if (v >= na) continue;
          if ((Ca[i] >> lane) & 1ull) {
            const u64 key = ((u64)(sm.deg[a_][v] + 1) << 32) | (u64)(0xFFFFFFFFu - v);
            bc = key > bc ? key : bc;
          }
          if ((Xa[i] >> lane) & 1ull) {
            const u64* row = rowsA + (u64)v * wb;
            bool elig = true;
            for (u32 w = 0; w < wb; ++w)
              if (sm.bad[a_ ^ 1][w] & ~row[w]) { elig = false; break; }
            if (elig) {
              const u64 key = ((u64)(sm.deg[a_][v] + 1) << 32) | (u64)(0xFFFFFFFFu - v);
              bx = key > bx ? key : bx;
            }
          }
        }
        bc = wave_max_u64(bc);
        bx = wave_max_u64(bx);
        // an X vertex wins only with a strictly larger degree
        if (bx != 0 && (bc == 0 || (bx >> 32) > (bc >> 32))) {
          dp[a_] = (int64_t)(bx >> 32) - 1; pv[a_] = 0xFFFFFFFFu - (u32)(bx & 0xFFFFFFFFu); pin[a_] = false;
        } else if (bc != 0) {
          dp[a_] = (int64_t)(bc >> 32) - 1; pv[a_] = 0xFFFFFFFFu - (u32)(bc & 0xFFFFFFFFu); pin[a_] = true;
        } else {
          dp[a_] = -1; pv[a_] = NONE; pin[a_] = false;
        }
      }
      pside = ((int64_t)nC[1] - dp[0] < (int64_t)nC[0] - dp[1]) ? 0 : 1;
      pivot = pv[pside];
      pivot_in_c = pin[pside];
      if (pivot != NONE) {
        const int os = pside ^ 1;
        const u32 w_os = os ? P.wr : P.wl;
        const u64* rowP = a.rows + (pside ? P.rows_r : P.rows_l) + (u64)pivot * w_os;
        const u64* Cos = setp(sm.set, os, 1, W);
        for (u32 w = lane; w < (u32)W; w += 64) sm.cand[w] = w < w_os ? (Cos[w] & ~rowP[w]) : 0;
        wave_barrier();
        kind = 2;
        nchild = popc_set<W>(sm.cand, w_os) + (pivot_in_c ? 1 : 0);
      }
    }
  }

finish:
  if (emit) emit_state<W>(sm, P, a.maps, a);
  if (nchild > 0) {
    u32 base = 0;
    if (lane == 0) base = atomicAdd(a.out_n, nchild);
    base = (u32)wave_bcast_u64(base, 0);
    const u32 out_cap = a.ctl ? a.ctl->out_cap : a.out_cap;
    if ((u64)base + nchild > out_cap) {
      // no room: mark the claimed slots below the cap as holes and requeue the
      // parent unchanged
      for (u32 s2 = base + lane; s2 < out_cap && s2 < base + nchild; s2 += 64)
        a.out[(u64)s2 * STRIDE] = INVALID_PID;
      // no room: requeue the parent unchanged
      u32 r = 0;
      if (lane == 0) r = atomicAdd(a.retry_n, 1u);
      r = (u32)wave_bcast_u64(r, 0);
      u64* o = a.retry + (u64)r * STRIDE;
      for (int j = lane; j < STRIDE; j += 64) o[j] = in[j];
      st.bytes += STRIDE * 8;
    } else if (kind == 1) {
      u64* o = a.out + (u64)base * STRIDE;
      write_child<W>(sm, P, a.rows, k, o, pid, bin_u, bin_side, bin_side, nullptr, NONE);
      wave_barrier();
      write_child<W>(sm, P, a.rows, k, o + STRIDE, pid, NONE, bin_side, bin_side, nullptr, bin_u);
      st.bytes += 2ull * STRIDE * 8;
    } else {
      const int os = pside ^ 1;
      const u32 w_os = os ? P.wr : P.wl;
      // running prefix of candidates already branched on (moved to X)
      u64* prefix = sm.prefix;
      for (int w = lane; w < W; w += 64) prefix[w] = 0;
      wave_barrier();
      u32 c = 0;
      for (u32 i = 0; i < w_os; ++i) {
        u64 s = sm.cand[i];
        while (s) {
          const u32 b = __ffsll((long long)s) - 1;
          s &= s - 1;
          const u32 u = i * 64 + b;
          write_child<W>(sm, P, a.rows, k, a.out + (u64)(base + c) * STRIDE, pid, u, os, os, prefix, NONE);
          ++c;
          wave_barrier();
          if (lane == 0) prefix[i] |= 1ull << b;
          wave_barrier();
        }
      }
      if (pivot_in_c) {
        write_child<W>(sm, P, a.rows, k, a.out + (u64)(base + c) * STRIDE, pid, pivot, pside, os, sm.cand, NONE);
        ++c;
      }
      st.bytes += (u64)c * STRIDE * 8;
    }
  }
}

template <int W>
__device__ void flush_stats(const ExpandArgs& a, const Stats& st, u32 stripe) {
  const u64 tsum = wave_sum_u64(st.intersections);
  const u64 bsum = wave_sum_u64(st.bytes);
  if ((threadIdx.x & 63) == 0) {
    u64* sp = a.stat + (stripe % NSTRIPE) * NSTAT;
    atomicAdd((unsigned long long*)&sp[0], (unsigned long long)st.states);
    atomicAdd((unsigned long long*)&sp[1], (unsigned long long)tsum);
    atomicAdd((unsigned long long*)&sp[3], (unsigned long long)bsum);
  }
}

// One block (one wave) per state.
template <int W>
__global__ __launch_bounds__(64) void bb_expand(ExpandArgs a) {
  __shared__ Smem<W> sm;
  if (blockIdx.x >= (a.ctl ? a.ctl->n_in : a.n_in)) return;
  if (__hip_atomic_load(a.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_LIMIT) return;
  Stats st = {0, 0, 0, 0};
  expand_one<W>(sm, a, blockIdx.x, st);
  flush_stats<W>(a, st, blockIdx.x);
}

// Persistent form: a fixed grid of waves takes states from the batch through
// a device counter until the batch is drained (or the result limit is hit),
// so uneven states balance across the chip without one block per state.
template <int W>
__global__ __launch_bounds__(64) void bb_expand_persistent(ExpandArgs a) {
  __shared__ Smem<W> sm;
  const int lane = threadIdx.x & 63;
  Stats st = {0, 0, 0, 0};
  const u32 n_in = a.ctl ? a.ctl->n_in : a.n_in;
  for (;;) {
    u32 idx = 0;
    if (lane == 0) idx = atomicAdd(a.work_next, 1u);
    idx = (u32)wave_bcast_u64(idx, 0);
    if (idx >= n_in) break;
    if (__hip_atomic_load(a.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_LIMIT) break;
    expand_one<W>(sm, a, idx, st);
  }
  flush_stats<W>(a, st, blockIdx.x);
}

// Wide partitions (W > W_LDS_MAX): the persistent form with each wave's
// working set in its own slice of a.scratch (HBM) instead of LDS.
template <int W>
__global__ __launch_bounds__(64) void bb_expand_wide(ExpandArgs a) {
  Smem<W>& sm = ((Smem<W>*)a.scratch)[blockIdx.x];
  const int lane = threadIdx.x & 63;
  Stats st = {0, 0, 0, 0};
  const u32 n_in = a.ctl ? a.ctl->n_in : a.n_in;
  for (;;) {
    u32 idx = 0;
    if (lane == 0) idx = atomicAdd(a.work_next, 1u);
    idx = (u32)wave_bcast_u64(idx, 0);
    if (idx >= n_in) break;
    if (__hip_atomic_load(a.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_LIMIT) break;
    expand_one<W>(sm, a, idx, st);
  }
  flush_stats<W>(a, st, blockIdx.x);
}

// Device-planned steps.  bb_plan runs one lane per width.  It first
// completes the width's last batch, whose retry list and children bb_commit
// has already copied back over the batch: the new top, the smoothed fan-out,
// and the check that some parent's children fitted.  With `plan` it then sets
// up the next batch as the host loop would (a batch of at most bmax states,
// fewer when the free room at the last fan-out is short; children may take
// half the stack) and clears the width's batch counters; a width whose stack
// is too full to expand safely halts (the host spills it), as does every width
// once the run hit its limit or overflowed its record buffer.
__global__ __launch_bounds__(64) void bb_plan(LaneCtl* ctl, u32* lane_cnt, const u32* flags, u32 active, int plan) {
  const u32 l = threadIdx.x;
  if (l >= 32 || !((active >> l) & 1u)) return;
  LaneCtl c = ctl[l];
  u32* cnt = lane_cnt + 4 * l;
  if (c.n_in) {
    const u64 nout = cnt[0] < c.out_cap ? cnt[0] : c.out_cap, nret = cnt[1];
    if (c.n_in > nret) c.fanout = 0.5f * c.fanout + 0.5f * (float)cnt[0] / (float)(c.n_in - nret);
    c.top = c.in_off + nret + nout;
    c.steps += 1;
    c.states += c.n_in;
    if (nret == c.n_in && cnt[0] >= c.out_cap) c.halt = 2;
    c.n_in = 0;
  }
  if (plan) {
    cnt[0] = cnt[1] = cnt[2] = 0;
    const u64 W = 1ull << l;
    const u64 room = c.cap - c.top;
    const u64 need = c.cap / 4 > 2 * (2 * 64 * W + 2) ? c.cap / 4 : 2 * (2 * 64 * W + 2);
    const u32 fl = __hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!c.halt && c.top > 0 && !(fl & (FLAG_LIMIT | FLAG_RES_OVERFLOW | FLAG_TABLE_FULL))) {
      if (c.top > 1 && room < need) {
        c.halt = 1;
      } else {
        const float f = c.fanout > 1.f ? c.fanout : 1.f;
        const u64 by_room = (u64)((double)room / (2.0 * (double)f));
        u64 B = c.top < c.bmax ? c.top : c.bmax;
        B = B < by_room ? B : by_room;
        B = B > 1 ? B : 1;
        const u64 oc = c.cap / 2 < room ? c.cap / 2 : room;
        if (oc == 0) {
          c.halt = 2;
        } else {
          c.n_in = (u32)B;
          c.in_off = c.top - B;
          c.out_cap = (u32)oc;
        }
      }
    }
  }
  ctl[l] = c;
}

// Copies a width's retry list and then its children (at most out_cap) over
// its batch on the stack, so the search stays depth-first across batches.
__global__ __launch_bounds__(256) void bb_commit(const LaneCtl* ctl, const u32* cnt, u64* stack, const u64* next,
                                                 const u64* retry, u32 stride) {
  const u32 n_in = ctl->n_in;
  if (n_in == 0) return;
  const u64 nret = cnt[1], nout = cnt[0] < ctl->out_cap ? cnt[0] : ctl->out_cap;
  const u64 wr = nret * stride, total = (nret + nout) * stride;
  u64* dst = stack + ctl->in_off * stride;
  for (u64 w = (u64)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (u64)gridDim.x * blockDim.x)
    dst[w] = w < wr ? retry[w] : next[w - wr];
}

}  // namespace gmk

namespace gmk {
hipError_t launch_bb_plan(LaneCtl* ctl, u32* lane_cnt, const u32* flags, u32 active, int plan, hipStream_t s) {
  hipLaunchKernelGGL(bb_plan, dim3(1), dim3(64), 0, s, ctl, lane_cnt, flags, active, plan);
  return hipGetLastError();
}

hipError_t launch_bb_commit(const LaneCtl* ctl, const u32* cnt, u64* stack, const u64* next, const u64* retry,
                            u32 stride, hipStream_t s) {
  hipLaunchKernelGGL(bb_commit, dim3(512), dim3(256), 0, s, ctl, cnt, stack, next, retry, stride);
  return hipGetLastError();
}

size_t bb_wide_scratch_bytes(int W) {
  switch (W) {
    case 256: return sizeof(Smem<256>);
    case 512: return sizeof(Smem<512>);
    case 1024: return sizeof(Smem<1024>);
    default: return 0;
  }
}

// Host-side launcher for the expansion kernel: one wave per state, or, with
// persistent_waves > 0, that many persistent waves draining the batch.  Widths
// above W_LDS_MAX always run persistent, with persistent_waves slices of
// a.scratch (bb_wide_scratch_bytes(W) each).
hipError_t launch_bb_expand(int W, u32 n_states, hipStream_t s, const ExpandArgs& a, u32 persistent_waves) {
  if (n_states == 0) return hipSuccess;
  dim3 block(64);
  if (W > W_LDS_MAX) {
    if (!persistent_waves || !a.scratch) return hipErrorInvalidValue;
    dim3 grid(n_states < persistent_waves ? n_states : persistent_waves);
    switch (W) {
      case 256: hipLaunchKernelGGL(bb_expand_wide<256>, grid, block, 0, s, a); break;
      case 512: hipLaunchKernelGGL(bb_expand_wide<512>, grid, block, 0, s, a); break;
      case 1024: hipLaunchKernelGGL(bb_expand_wide<1024>, grid, block, 0, s, a); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (persistent_waves) {
    dim3 grid(n_states < persistent_waves ? n_states : persistent_waves);
    switch (W) {
      case 1: hipLaunchKernelGGL(bb_expand_persistent<1>, grid, block, 0, s, a); break;
      case 2: hipLaunchKernelGGL(bb_expand_persistent<2>, grid, block, 0, s, a); break;
      case 4: hipLaunchKernelGGL(bb_expand_persistent<4>, grid, block, 0, s, a); break;
      case 8: hipLaunchKernelGGL(bb_expand_persistent<8>, grid, block, 0, s, a); break;
      case 16: hipLaunchKernelGGL(bb_expand_persistent<16>, grid, block, 0, s, a); break;
      case 32: hipLaunchKernelGGL(bb_expand_persistent<32>, grid, block, 0, s, a); break;
      case 64: hipLaunchKernelGGL(bb_expand_persistent<64>, grid, block, 0, s, a); break;
      case 128: hipLaunchKernelGGL(bb_expand_persistent<128>, grid, block, 0, s, a); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  dim3 grid(n_states);
  switch (W) {
    case 1: hipLaunchKernelGGL(bb_expand<1>, grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL(bb_expand<2>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(bb_expand<4>, grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL(bb_expand<8>, grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL(bb_expand<16>, grid, block, 0, s, a); break;
    case 32: hipLaunchKernelGGL(bb_expand<32>, grid, block, 0, s, a); break;
    case 64: hipLaunchKernelGGL(bb_expand<64>, grid, block, 0, s, a); break;
    case 128: hipLaunchKernelGGL(bb_expand<128>, grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
}  // namespace gmk
