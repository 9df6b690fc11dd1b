// Thin-result kernels (see thin.h): family construction (a lane per tuple, or
// a wave per tuple for one-vertex tuples in small launches and for tuples of
// more than k vertices), the per-launch rank index, and concrete-combination
// checking (a lane per rank, then a wave per rank for those with long
// candidate lists) with emission a wave at a time; also the neighbour-hash
// build used by the checks' adjacency tests.
#include "gmk_device.h"
#include "thin.h"
#include "thin_build.h"
#include "thin_args.h"

namespace gmk {

struct DevArena {
  u32* base;
  u64 cap;
  u64* used;
  __device__ u64 alloc(u64 n) {
    const u64 at = atomicAdd((unsigned long long*)used, (unsigned long long)n);
    return at + n <= cap ? at : thin::SAT;
  }
  __device__ u32* ptr(u64 off) { return base + off; }
};

struct DevFamilies {
  thin::Family* fams;
  u32 cap;
  u32* n;
  u32* flags;
  __device__ thin::Family* next() {
    const u32 i = atomicAdd(n, 1u);
    if (i >= cap) { atomicOr(flags, FLAG_TABLE_FULL); return nullptr; }
    return fams + i;
  }
};

// One lane per tuple: one-vertex tuples (O(1): the lists are the CSR row) and
// tuples whose far pool needs N(L) in ascending order (sequential merge).  J is
// the tuple size (j = 1, or j <= k <= 3), so the per-lane type arrays have 2^J
// entries.
template <int J>
__global__ __launch_bounds__(256) void thin_build_families(ThinArgs a) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  thin::Cost cost;
  if (t < a.ntuples) {
    DevArena ar{a.arena, a.arena_cap, a.arena_used};
    DevFamilies fo{a.fams, a.fam_cap, a.fam_n, a.tflags};
    if (!thin::build_families_t<J>(a.g, a.tuples + (u64)t * J, J, a.k, a.min_b, t, ar, fo, cost))
      atomicOr(a.tflags, FLAG_RES_OVERFLOW);
  }
  const u64 b = wave_sum_u64(cost.bytes);
  if ((threadIdx.x & 63) == 0 && b)
    atomicAdd((unsigned long long*)&a.stat[((t >> 6) % NSTRIPE) * NSTAT + 4], (unsigned long long)b);
}

// The k+1 free vertices of least degree (ties: first in the list), as
// free_picks, by a whole wave: each lane keeps its best k+1 of a strided share,
// then k+1 rounds of a wave minimum take the overall best.  Returns their
// number.
__device__ __forceinline__ u32 wave_free_picks(const ThinArgs& a, const u32* freel, u32 nfree, u32 (&fpick)[4],
                                               thin::Cost& cost) {
  const int lane = threadIdx.x & 63;
  const u32 npk = a.k + 1 < 4 ? a.k + 1 : 4;
  u64 best[4] = {~0ull, ~0ull, ~0ull, ~0ull};
  for (u32 i = lane; i < nfree; i += 64) {
    const u32 y = freel[i];
    u64 key = ((u64)(a.g.off[y + 1] - a.g.off[y]) << 32) | i;
    cost.bytes += 12;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if ((u32)q < npk && key < best[q]) { const u64 tmp = best[q]; best[q] = key; key = tmp; }
  }
  u32 nfp = 0, head = 0;
  for (u32 q = 0; q < npk; ++q) {
    u64 mine = ~0ull;
#pragma unroll
    for (int h = 0; h < 4; ++h)
      if ((u32)h == head) mine = best[h];
    const u64 m = wave_min_u64(mine);
    if (m == ~0ull) break;
    if (mine == m) ++head;
    const u32 y = freel[(u32)(m & 0xFFFFFFFFu)];
#pragma unroll
    for (int h = 0; h < 4; ++h)
      if ((u32)h == nfp) fpick[h] = y;
    ++nfp;
  }
  return nfp;
}

// Family slots of the ordered build: counting only, or consecutive slots from
// a tuple's offset.
struct CountFamilies {
  u32 n = 0;
  __device__ thin::Family* next() { ++n; return nullptr; }
};
struct SlotFamilies {
  thin::Family* at;
  __device__ thin::Family* next() { return at++; }
};

// One wave per one-vertex tuple (launches too small to fill the device one
// lane per tuple): as build_families_t for j = 1 (free list = the vertex's CSR
// row, far pool = everything else), with the lanes splitting the witness scan
// and the choice of the k+1 lightest free vertices.  Families are written in
// tuple order, as the lane-per-tuple form of a small launch does, since the
// order decides which ranks a result limit reaches first: this kernel counts
// each tuple's families and keeps its witness bound and lightest vertices,
// thin_scan_counts turns the counts into offsets, and thin_emit_wave1 writes
// the families there.
//   tcount[t]: families of tuple t; tcount[n + t]: their offset;
//   tcount[2n + 6t ..]: witness bound, number of picks, up to 4 picks.
__global__ __launch_bounds__(256) void thin_build_wave1(ThinArgs a) {
  const int lane = threadIdx.x & 63;
  const u32 t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= a.ntuples) return;  // uniform per wave
  const u32* L = a.tuples + t;
  thin::Cost cost;
  const u32 s = L[0];
  const u32* ns = a.g.adj + a.g.off[s];
  const u32 d = a.g.off[s + 1] - a.g.off[s];
  cost.bytes += 8;
  const u32 dom = (u32)wave_min_u64((u64)thin::dominance_bound(a.g, L, 1, a.k, ns, d, (u32)lane, 64u, cost));
  u32 fpick[4] = {0, 0, 0, 0};
  const u32 nfp = wave_free_picks(a, ns, d, fpick, cost);
  if (lane == 0) {
    u32 mcnt[2] = {d, 0};
    u64 poolp[2] = {(u64)ns, 0};
    CountFamilies fo;
    thin::Cost none;
    thin::emit_families<1>(L, 1, a.k, a.min_b, t, mcnt, (u64)(a.g.b_hi - a.g.b_lo) - d, poolp, ns, d, dom, fpick,
                           nfp, fo, none);
    a.tcount[t] = fo.n;
    u32* keep = a.tcount + 2ull * a.ntuples + 6ull * t;
    keep[0] = dom;
    keep[1] = nfp;
#pragma unroll
    for (int q = 0; q < 4; ++q) keep[2 + q] = fpick[q];
  }
  const u64 b = wave_sum_u64(cost.bytes);
  if (lane == 0 && b) atomicAdd((unsigned long long*)&a.stat[(t % NSTRIPE) * NSTAT + 4], (unsigned long long)b);
}

// One lane per one-vertex tuple: its families, from what thin_build_wave1
// kept, at the offset thin_scan_counts gave it.
__global__ __launch_bounds__(256) void thin_emit_wave1(ThinArgs a) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  thin::Cost cost;
  if (t < a.ntuples && __hip_atomic_load(a.fam_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= a.fam_cap) {
    const u32* L = a.tuples + t;
    const u32 s = L[0];
    const u32* ns = a.g.adj + a.g.off[s];
    const u32 d = a.g.off[s + 1] - a.g.off[s];
    const u32* keep = a.tcount + 2ull * a.ntuples + 6ull * t;
    const u32 fpick[4] = {keep[2], keep[3], keep[4], keep[5]};
    u32 mcnt[2] = {d, 0};
    u64 poolp[2] = {(u64)ns, 0};
    SlotFamilies fo{a.fams + a.tcount[a.ntuples + t]};
    cost.bytes += 8 + 24;
    thin::emit_families<1>(L, 1, a.k, a.min_b, t, mcnt, (u64)(a.g.b_hi - a.g.b_lo) - d, poolp, ns, d, keep[0], fpick,
                           keep[1], fo, cost);
  }
  const u64 b = wave_sum_u64(cost.bytes);
  if ((threadIdx.x & 63) == 0 && b)
    atomicAdd((unsigned long long*)&a.stat[((t >> 6) % NSTRIPE) * NSTAT + 4], (unsigned long long)b);
}

// One wave per tuple of j >= 2 vertices without a far pool (j > k): lanes sweep
// the members' neighbour lists in coalesced chunks; an element is counted once,
// by the first list that holds it, with its miss type found by binary search in
// the other lists.  Pass 1 counts the types, pass 2 writes the pools (the free
// list in ascending order, other pools in any order); lane 0 then emits the
// families.
template <int J>
__global__ __launch_bounds__(256) void thin_build_wave(ThinArgs a) {
  __shared__ u32 s_off[4][1 << thin::MAXJ];  // pass 1: counts; pass 2: pool offsets
  __shared__ u32 s_cur[4][1 << thin::MAXJ];  // pass 2: fill cursors
  __shared__ u64 s_base[4];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const u32 t = blockIdx.x * 4 + wv;
  if (t >= a.ntuples) return;  // uniform per wave
  constexpr u32 j = J, full = (1u << J) - 1u;
  const u32* L = a.tuples + (u64)t * j;
  thin::Cost cost;
  u32 lo[J], dg[J];
  for (u32 i = 0; i < j; ++i) { lo[i] = a.g.off[L[i]]; dg[i] = a.g.off[L[i] + 1] - lo[i]; }
  s_off[wv][lane] = 0;
  s_cur[wv][lane] = 0;
  wave_barrier();
  u32 cnt_lane = 0;  // lane m keeps the count of type m
  for (int pass = 0; pass < 2; ++pass) {
    u32* pools = pass ? a.arena + s_base[wv] : nullptr;
    u32 free_fill = 0;
    for (u32 i = 0; i < j; ++i) {
      const u32* li = a.g.adj + lo[i];
      for (u32 base = 0; base < dg[i]; base += 64) {
        const u32 e = base + lane;
        const bool valid = e < dg[i];
        u32 hit = 1u << i;
        bool first = valid;
        u32 y = 0;
        if (valid) {
          y = li[e];
          cost.bytes += 4;
          for (u32 q = 0; q < j; ++q) {
            if (q == i) continue;
            if (thin::contains(a.g.adj + lo[q], dg[q], y, cost)) {
              hit |= 1u << q;
              if (q < i) first = false;
            }
          }
        }
        const u32 m = full & ~hit;
        if (pass == 0) {
          if (first) atomicAdd(&s_off[wv][m], 1u);
        } else {
          // free vertices come only from list 0: ordered compaction keeps them ascending
          const u64 fb = __ballot(first && m == 0);
          if (first) {
            const u32 at = m == 0 ? free_fill + (u32)__popcll(fb & ((1ull << lane) - 1ull))
                                  : s_off[wv][m] + atomicAdd(&s_cur[wv][m], 1u);
            pools[at] = y;
            cost.bytes += 4;
          }
          free_fill += (u32)__popcll(fb);
        }
      }
    }
    wave_barrier();
    if (pass == 1) break;
    cnt_lane = s_off[wv][lane];
    if (lane == 0) {
      u32 d = 0;
      u64 maxr = s_off[wv][0];
      for (u32 mm = 0; mm <= full; ++mm) {
        const u32 cm = s_off[wv][mm];
        d += cm;
        if (mm > 0 && mm < full && (u32)thin::popc(mm) <= a.k) maxr += cm < a.k ? cm : a.k;
      }
      u64 at = thin::SAT;
      if (maxr >= a.min_b && maxr > 0) {
        DevArena ar{a.arena, a.arena_cap, a.arena_used};
        at = ar.alloc(d);
        if (at == thin::SAT) atomicOr(a.tflags, FLAG_RES_OVERFLOW);
      }
      s_base[wv] = at;
      u32 acc = 0;
      for (u32 mm = 0; mm <= full; ++mm) { const u32 cm = s_off[wv][mm]; s_off[wv][mm] = acc; acc += cm; }
    }
    wave_barrier();
    if (s_base[wv] == thin::SAT) break;
  }
  if (s_base[wv] != thin::SAT) {
    u32 mcnt[1 << J];
    u32 d = 0;
    for (u32 mm = 0; mm <= full; ++mm) { mcnt[mm] = (u32)__shfl((int)cnt_lane, (int)mm, 64); d += mcnt[mm]; }
    // the free list (type 0) starts the tuple's pools; lanes split the witness scan
    const u32* freel = a.arena + s_base[wv] + s_off[wv][0];
    u32 dom = thin::dominance_bound(a.g, L, j, a.k, freel, mcnt[0], (u32)lane, 64u, cost);
    dom = (u32)wave_min_u64((u64)dom);
    u32 fpick[4];
    const u32 nfp = wave_free_picks(a, freel, mcnt[0], fpick, cost);
    if (lane == 0) {
      u64 poolp[1 << J];
      for (u32 mm = 0; mm <= full; ++mm) poolp[mm] = (u64)(a.arena + s_base[wv] + s_off[wv][mm]);
      DevFamilies fo{a.fams, a.fam_cap, a.fam_n, a.tflags};
      thin::emit_families<J>(L, j, a.k, a.min_b, t, mcnt, (u64)(a.g.b_hi - a.g.b_lo) - d, poolp, nullptr, d, dom,
                             fpick, nfp, fo, cost);
    }
  }
  const u64 b = wave_sum_u64(cost.bytes);
  if (lane == 0 && b) atomicAdd((unsigned long long*)&a.stat[(t % NSTRIPE) * NSTAT + 4], (unsigned long long)b);
}

// Exclusive prefix sums of the per-tuple family counts cnt[0..n) into
// cnt[n..2n), one block of 1024 threads; sets the launch's family count and
// flags a full table.
__global__ __launch_bounds__(1024) void thin_scan_counts(u32* cnt, u32 n, u32* fam_n, u32 cap, u32* tflags) {
  __shared__ u32 wsum[16];
  const u32 t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const u32 seg = (n + 1023) / 1024;
  const u32 b = t * seg < n ? t * seg : n;
  const u32 e = b + seg < n ? b + seg : n;
  u32 s = 0;
  for (u32 i = b; i < e; ++i) s += cnt[i];
  u32 inc = s;
  for (int d = 1; d < 64; d <<= 1) {
    const u32 o = __shfl_up(inc, d, 64);
    if ((int)lane >= d) inc += o;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  if (t == 0) {
    u32 run = 0;
    for (int w = 0; w < 16; ++w) { const u32 v = wsum[w]; wsum[w] = run; run += v; }
  }
  __syncthreads();
  u32 run = wsum[wv] + inc - s;
  for (u32 i = b; i < e; ++i) { cnt[n + i] = run; run += cnt[i]; }
  if (t == 1023) {
    *fam_n = run;
    if (run > cap) atomicOr(tflags, FLAG_TABLE_FULL);
  }
}

// Exclusive prefix sums of the family totals (saturating at SAT), one block of
// TSCAN threads: per-thread segment sums, an inclusive scan of those in LDS,
// then each thread's segment from the sum of the segments before it (the
// phases are thin::scan_phase_*).  A family whose prefix is finite keeps it
// even when the running sum saturates inside its segment.
// With nfam_p the family count is read on the device (at most fam_cap), and
// total_out receives the sum.
__global__ __launch_bounds__(thin::TSCAN) void thin_scan_totals(const thin::Family* fams, u32 nfam, const u32* nfam_p,
                                                                 u32 fam_cap, u64* prefix, u64 base, u64 cap,
                                                                 u64* total_out) {
  __shared__ u64 part[thin::TSCAN];
  const u32 t = threadIdx.x;
  if (nfam_p) nfam = *nfam_p < fam_cap ? *nfam_p : fam_cap;
  auto total = [fams, base, cap](u32 i) {
    const u64 v = fams[i].total, rest = v > base ? v - base : 0;
    return rest < cap ? rest : cap;
  };
  part[t] = thin::scan_phase_sum(nfam, t, total);
  __syncthreads();
  for (u32 d = 1; d < thin::TSCAN; d <<= 1) {
    const u64 v = thin::scan_phase_step(part, t, d);
    __syncthreads();
    part[t] = v;
    __syncthreads();
  }
  thin::scan_phase_write(nfam, t, part, total, prefix);
  if (total_out && t == thin::TSCAN - 1) *total_out = part[thin::TSCAN - 1];
}

template <int K>
__device__ __forceinline__ void flush_check_stats(const ThinArgs& a, const thin::Cost& cost, u64 n_emit,
                                                  u64 n_tested) {
  const u64 se = wave_sum_u64(n_emit), stt = wave_sum_u64(n_tested), sb = wave_sum_u64(cost.bytes);
  if ((threadIdx.x & 63) == 0 && stt) {
    u64* sp = a.stat + ((blockIdx.x * 4 + (threadIdx.x >> 6)) % NSTRIPE) * NSTAT;
    atomicAdd((unsigned long long*)&sp[1], (unsigned long long)stt);
    atomicAdd((unsigned long long*)&sp[2], (unsigned long long)se);
    atomicAdd((unsigned long long*)&sp[5], (unsigned long long)sb);
  }
}

// Probe runs: adds a rank's outcome and the lane time since t0 to its tuple's
// log entry.
__device__ __forceinline__ void log_item(const ThinArgs& a, const thin::Family* f, bool emitted, u64 t0) {
  if (!a.item_log || !f) return;
  const u64 dt = (u64)wall_clock64() - t0;
  if (emitted) atomicAdd((unsigned long long*)&a.item_log[2 * f->tuple], 1ull);
  atomicAdd((unsigned long long*)&a.item_log[2 * f->tuple + 1], (unsigned long long)dt);
}

// K: misses allowed per vertex (1..3).  A family chooses at most 2 K^2 vertices
// besides its free ones (each costs budget of at least one of <= 2K members),
// so the chosen vertices live in registers: every access uses a constant index.

// Family of rank r and its chosen vertices (with their type sizes); null when r
// lies past the end of the family the prefix sums give (ranks at or above SAT
// when the sums saturate are never launched, so this does not happen; the
// check keeps a wrong prefix from decoding an out-of-range offset).
template <int K>
__device__ __forceinline__ const thin::Family* unrank_rank(const ThinArgs& a, u64 r, u32 (&ch)[2 * K * K],
                                                          u32 (&ctsz)[2 * K * K], u32& nch, thin::Cost& cost) {
  constexpr int NM = 2 * K * K;
#pragma unroll
  for (int q = 0; q < NM; ++q) { ch[q] = 0; ctsz[q] = 0; }
  nch = 0;
  // the families of the first ranks of r's block and of the next block bound r's
  const u64 b = (r - a.r0) >> RIDX_LOG;
  u32 lo = a.ridx[b], hi = a.ridx[b + 1] + 1;
  cost.bytes += 8;
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (a.prefix[mid] <= r) lo = mid; else hi = mid;
  }
  const thin::Family* f = a.fams + lo;
  u64 rr = r - a.prefix[lo] + a.rank_base;
  cost.bytes += 8ull * 8 + 16ull * f->ntypes;  // prefix probes, family header and type fields
  if (rr >= f->total) return nullptr;  // past the family's combinations: no rank
  const u32 nt = f->ntypes;
  for (u32 t = 0; t < nt; ++t) {
    const u32 cnt = f->cnt[t], ps = f->psize[t];
    // mixed radix over the types; the last digit is what remains (rr < its radix)
    u64 rt = rr;
    if (t + 1 < nt) {
      const u64 radix = thin::binom(ps, cnt);
      rt = rr % radix;
      rr /= radix;
    }
    u32 idx[3] = {0, 0, 0};
    thin::unrank(rt, cnt, ps, idx);
    const u32 tsz = thin::popc(f->mask[t]);
    // a run-time loop over the chosen indices keeps one copy of the pool lookup
    for (u32 q = 0; q < cnt; ++q) {
      const u32 iq = q == 0 ? idx[0] : (q == 1 ? idx[1] : idx[2]);
      const u32 y = thin::pool_member(*f, t, iq, a.g.b_lo, cost);
#pragma unroll
      for (int s2 = 0; s2 < NM; ++s2)
        if ((u32)s2 == nch) { ch[s2] = y; ctsz[s2] = tsz; }
      ++nch;
    }
  }
  return f;
}

// What the checks read of family f: its guard table is used for whole-family
// decisions (no candidate) always, and for candidate lists when the launch
// allows them (ThinArgs.guard_lists).
__device__ __forceinline__ thin::FamRef fam_view(const ThinArgs& a, const thin::Family& f) {
  thin::FamRef r = thin::fam_ref(f);
  if (!a.guard_lists && r.gmode == thin::GUARD_TABLE) r.gmode = thin::GUARD_NONE;
  return r;
}

// Wave-wide reductions by DPP moves within each row of 16 lanes (quad_perm
// [1,0,3,2] and [2,3,0,1], then row_ror 4 and 8: every lane ends with its row's
// total) and a readlane of the four row totals: ALU work only, no LDS
// crossbar traffic.  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ u32 dpp_u32(u32 v) {
  return (u32)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_ROR4 = 0x124, DPP_ROR8 = 0x128;
__device__ __forceinline__ u32 rows_u32(u32 v, int r) { return (u32)__builtin_amdgcn_readlane((int)v, 16 * r); }
__device__ __forceinline__ u32 wave_sum_u32(u32 v) {
  v += dpp_u32<DPP_XOR1>(v);
  v += dpp_u32<DPP_XOR2>(v);
  v += dpp_u32<DPP_ROR4>(v);
  v += dpp_u32<DPP_ROR8>(v);
  return rows_u32(v, 0) + rows_u32(v, 1) + rows_u32(v, 2) + rows_u32(v, 3);
}
template <int CTRL>
__device__ __forceinline__ u64 dpp_u64(u64 v) {
  return ((u64)dpp_u32<CTRL>((u32)(v >> 32)) << 32) | dpp_u32<CTRL>((u32)v);
}
__device__ __forceinline__ u64 rows_u64(u64 v, int r) {
  return ((u64)rows_u32((u32)(v >> 32), r) << 32) | rows_u32((u32)v, r);
}
__device__ __forceinline__ u64 wave_sum_u64_dpp(u64 v) {
  v += dpp_u64<DPP_XOR1>(v);
  v += dpp_u64<DPP_XOR2>(v);
  v += dpp_u64<DPP_ROR4>(v);
  v += dpp_u64<DPP_ROR8>(v);
  return rows_u64(v, 0) + rows_u64(v, 1) + rows_u64(v, 2) + rows_u64(v, 3);
}
__device__ __forceinline__ u64 wave_xor_u64_dpp(u64 v) {
  v ^= dpp_u64<DPP_XOR1>(v);
  v ^= dpp_u64<DPP_XOR2>(v);
  v ^= dpp_u64<DPP_ROR4>(v);
  v ^= dpp_u64<DPP_ROR8>(v);
  return rows_u64(v, 0) ^ rows_u64(v, 1) ^ rows_u64(v, 2) ^ rows_u64(v, 3);
}

// Exclusive prefix sums over the wave.
__device__ __forceinline__ u32 wave_excl_scan_u32(u32 v) {
  const int lane = threadIdx.x & 63;
  u32 inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const u32 o = (u32)__shfl_up((int)inc, (unsigned)d, 64);
    if (lane >= d) inc += o;
  }
  return inc - v;
}

__device__ __forceinline__ u64 wave_excl_scan_u64(u64 v) {
  const int lane = threadIdx.x & 63;
  u64 inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const u64 o = wave_bcast_u64(inc, lane >= d ? lane - d : lane);
    if (lane >= d) inc += o;
  }
  return inc - v;
}

// Has the result limit been reached?  The same answer in every lane.
__device__ __forceinline__ bool limit_reached(const ThinArgs& a) {
  if (a.unlimited) return false;
  return __ballot((__hip_atomic_load(a.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_LIMIT) != 0) != 0;
}

// Fingerprint of the maximal (L, R) of family f with chosen vertices ch, and
// its dedup insert when the table is on: returns whether the result stays
// (false: a repeat the table already holds).
template <int K>
__device__ __forceinline__ bool lane_fingerprint(const ThinArgs& a, const thin::Family* f, const u32 (&ch)[2 * K * K],
                                                 u32 nch, u32 stripe, u64& fpv, thin::Cost& cost) {
  constexpr int NM = 2 * K * K;
  const u32 nr0 = f->nfree + nch;
  const u32* fr0 = f->freep;
  u64 h = 0;
  for (u32 i = 0; i < f->j; ++i) h += mix64(f->L[i]);
  for (u32 i = 0; i < f->nfree; ++i) h += mix64(fr0[i]);
#pragma unroll
  for (int i = 0; i < NM; ++i)
    if ((u32)i < nch) h += mix64(ch[i]);
  cost.bytes += 4ull * f->nfree;
  const u64 nlr = a.swap_sides ? nr0 : f->j, nrr = a.swap_sides ? f->j : nr0;
  fpv = fmix64(h + nlr * FP_A + nrr * FP_B);
  if (!a.dedup.on) return true;
  // the table's check word: an independent hash of the same record
  u64 h2 = 0;
  for (u32 i = 0; i < f->j; ++i) h2 += mix64b(f->L[i]);
  for (u32 i = 0; i < f->nfree; ++i) h2 += mix64b(fr0[i]);
#pragma unroll
  for (int i = 0; i < NM; ++i)
    if ((u32)i < nch) h2 += mix64b(ch[i]);
  cost.bytes += 4ull * f->nfree;
  return dedup_insert(a.dedup, fpv, fmix64(h2 + nlr * FP_B + nrr * FP_A), a.flags, stripe);
}

// Writes the record [|L|, L..., |R|, R...] (left side first) of family f with
// chosen vertices ch at out (device or LDS memory); R merges the free list
// with the chosen vertices, sorted in place (when they are not ascending
// already) by an odd-even transposition network over the first nch slots.
// ch is consumed.
template <int K>
__device__ __forceinline__ void write_record(u32* out, const ThinArgs& a, const thin::Family* f, u32 (&ch)[2 * K * K],
                                             u32 nch) {
  constexpr int NM = 2 * K * K;
  const u32 nl = f->j, nr = f->nfree + nch;
  // chosen vertices of one type come ascending (the far pool always, as the
  // lists of one-vertex tuples): the network runs only when they are not
  bool ordered = true;
#pragma unroll
  for (int q = 0; q + 1 < NM; ++q) ordered = ordered && ((u32)(q + 1) >= nch || ch[q] <= ch[q + 1]);
  if (!ordered) {
#pragma unroll
    for (int pass = 0; pass < NM; ++pass)
#pragma unroll
      for (int q = pass & 1; q + 1 < NM; q += 2)
        if ((u32)(q + 1) < nch && ch[q] > ch[q + 1]) { const u32 tmp = ch[q]; ch[q] = ch[q + 1]; ch[q + 1] = tmp; }
  }
  const u32* fr = f->freep;
  // tuple side and generated side, left side first
  const u32 at_t = a.swap_sides ? 1 + nr + 1 : 1;
  const u32 at_r = a.swap_sides ? 1 : 1 + nl + 1;
  out[a.swap_sides ? 1 + nr : 0] = nl;
  out[a.swap_sides ? 0 : 1 + nl] = nr;
  for (u32 i = 0; i < nl; ++i) out[at_t + i] = f->L[i];
  u32 p = 0, q = 0;
  if constexpr (K <= 2) {
    // merge: the next chosen vertex is always ch[0]; taking it shifts the
    // rest down one slot (constant indices, so ch stays in registers)
    u32 cq = nch ? ch[0] : 0xFFFFFFFFu;
    for (u32 i = 0; i < nr; ++i) {
      u32 v;
      if (q >= nch || (p < f->nfree && fr[p] < cq)) {
        v = fr[p++];
      } else {
        v = cq;
        ++q;
#pragma unroll
        for (int s2 = 0; s2 + 1 < NM; ++s2) ch[s2] = ch[s2 + 1];
        cq = q < nch ? ch[0] : 0xFFFFFFFFu;
      }
      out[at_r + i] = v;
    }
  } else {
    // k = 3 (18 slots): the shifts would cost a wave per SIMD in registers;
    // ch[q] is fetched by a select over constant indices instead
    for (u32 i = 0; i < nr; ++i) {
      u32 cq = 0xFFFFFFFFu;  // ch[q], or none
#pragma unroll
      for (int s2 = 0; s2 < NM; ++s2)
        if ((u32)s2 == q && q < nch) cq = ch[s2];
      u32 v;
      if (q >= nch || (p < f->nfree && fr[p] < cq)) v = fr[p++];
      else { v = cq; ++q; }
      out[at_r + i] = v;
    }
  }
}

// Emits the maximal (L, R) of family f with chosen vertices ch for every lane
// with `want` set: fingerprint, dedup, result slot under the limit, record.
// Every lane of the wave calls it together: the wave takes its result slots,
// record space and fingerprint sums with one atomic each instead of one per
// result, since results from every wave meet on the same counters.  Returns
// whether this lane's result was emitted.
template <int K>
__device__ __forceinline__ bool emit_wave(const ThinArgs& a, bool want, const thin::Family* f, u32 (&ch)[2 * K * K],
                                          u32 nch, u32 stripe, thin::Cost& cost) {
  const int lane = threadIdx.x & 63;
  u64 fpv = 0;
  if (want) want = lane_fingerprint<K>(a, f, ch, nch, stripe, fpv, cost);
  const u64 mask = __ballot(want);
  if (!mask) return false;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  const u32 nwant = (u32)__popcll(mask);
  u64 base = 0;
  if (lane == leader) {
    base = atomicAdd((unsigned long long*)a.res_count, (unsigned long long)nwant);
    if (base + nwant >= a.limit) atomicOr(a.flags, FLAG_LIMIT);
  }
  base = wave_bcast_u64(base, leader);
  const u64 slot = base + (u64)__popcll(mask & ((1ull << lane) - 1));
  if (want && slot >= a.limit) want = false;
  // fingerprint sums of the emitted results
  const u64 fsum = wave_sum_u64(want ? fpv : 0);
  u64 fxor = want ? fpv : 0;
  for (int m = 32; m > 0; m >>= 1) fxor ^= shfl_xor_u64(fxor, m);
  if (lane == leader && fsum | fxor) {
    u64* fs = a.fp + (stripe % NSTRIPE) * 2;
    atomicAdd((unsigned long long*)&fs[0], (unsigned long long)fsum);
    atomicXor((unsigned long long*)&fs[1], (unsigned long long)fxor);
  }
  if (!a.keep) return want;
  const u32 nl = f ? f->j : 0, nr = want ? f->nfree + nch : 0;
  const u64 len = want ? 2ull + nl + nr : 0;
  const u64 off = wave_excl_scan_u64(len);
  const u64 tot = wave_bcast_u64(off + len, 63);
  if (!tot) return want;
  u64 pos = 0;
  if (lane == leader) pos = atomicAdd((unsigned long long*)a.res_words, (unsigned long long)tot);
  pos = wave_bcast_u64(pos, leader) + off;
  if (!want) return false;
  cost.bytes += 4ull * len;
  if (pos + len > a.res_cap) {
    atomicOr(a.flags, FLAG_RES_OVERFLOW);
    return true;
  }
  write_record<K>(a.res + pos, a, f, ch, nch);
  return true;
}

// Waves per workgroup of the lane-per-rank check, and the record words a
// workgroup stages in LDS before writing them out with coalesced stores (more
// go straight to device memory).
constexpr u32 CHECK_WAVES = 4;
constexpr u32 CHECK_THREADS = 64 * CHECK_WAVES;
constexpr u32 STAGE_WORDS = 1024 * CHECK_WAVES;

// emit_wave for a whole workgroup of 4 waves (every thread calls it; used by
// the lane-per-rank check, whose workgroups reach it together, each lane with
// at most one tested rank): result slots, record space, fingerprint sums and
// the check statistics are taken with one atomic each per workgroup, and the
// workgroup's records, which fill one contiguous range, are assembled in LDS
// and written with coalesced stores.  Counts come from ballots, sums from DPP
// reductions, and the record offsets from a 32-bit scan.
template <int K>
__device__ __forceinline__ bool emit_block(const ThinArgs& a, bool want, const thin::Family* f, u32 (&ch)[2 * K * K],
                                           u32 nch, bool tested, thin::Cost& cost) {
  __shared__ u64 s_n[CHECK_WAVES], s_w[CHECK_WAVES], s_fs[CHECK_WAVES], s_fx[CHECK_WAVES], s_st[CHECK_WAVES][3];
  __shared__ u64 s_base, s_wbase;
  __shared__ u32 s_stage[STAGE_WORDS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const u32 stripe = blockIdx.x;
  u64 fpv = 0;
  if (want) want = lane_fingerprint<K>(a, f, ch, nch, stripe, fpv, cost);
  const u64 mask = __ballot(want);
  if (lane == 0) s_n[wv] = (u64)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 tot = 0;
#pragma unroll
    for (u32 w = 0; w < CHECK_WAVES; ++w) tot += s_n[w];
    u64 base = 0;
    if (tot && a.unlimited) {
      // every result of the launch fits under the limit: the count is added
      // without waiting for the old value
      atomicAdd((unsigned long long*)a.res_count, (unsigned long long)tot);
    } else if (tot) {
      base = atomicAdd((unsigned long long*)a.res_count, (unsigned long long)tot);
      if (base + tot >= a.limit) atomicOr(a.flags, FLAG_LIMIT);
    }
    s_base = base;
  }
  __syncthreads();
  u64 before = 0;
#pragma unroll
  for (int w = 0; w < (int)CHECK_WAVES; ++w) before += w < wv ? s_n[w] : 0ull;
  const u64 slot = s_base + before + (u64)__popcll(mask & ((1ull << lane) - 1));
  if (want && !a.unlimited && slot >= a.limit) want = false;
  const u32 nl = want ? f->j : 0, nr = want ? f->nfree + nch : 0;
  const u32 len = want && a.keep ? 2u + nl + nr : 0u;
  cost.bytes += 4ull * len;
  const u32 off = wave_excl_scan_u32(len);
  const u32 wtot = (u32)__builtin_amdgcn_readlane((int)(off + len), 63);
  const u64 emitted = __ballot(want), tested_mask = __ballot(tested);
  const u64 fsum = wave_sum_u64_dpp(want ? fpv : 0);
  const u64 fxor = wave_xor_u64_dpp(want ? fpv : 0);
  // a lane's bytes for one rank fit 32 bits
  const u32 sb = wave_sum_u32(cost.bytes < 0xFFFFFFFFull ? (u32)cost.bytes : 0xFFFFFFFFu);
  if (lane == 0) {
    s_w[wv] = wtot;
    s_fs[wv] = fsum;
    s_fx[wv] = fxor;
    s_st[wv][0] = (u64)__popcll(tested_mask);
    s_st[wv][1] = (u64)__popcll(emitted);
    s_st[wv][2] = sb;
  }
  __syncthreads();
  u64 btot = 0;
#pragma unroll
  for (u32 w = 0; w < CHECK_WAVES; ++w) btot += s_w[w];
  if (threadIdx.x == 0) {
    u64 fs = 0, fx = 0, tt = 0, te = 0, tb = 0;
#pragma unroll
    for (u32 w = 0; w < CHECK_WAVES; ++w) {
      fs += s_fs[w];
      fx ^= s_fx[w];
      tt += s_st[w][0];
      te += s_st[w][1];
      tb += s_st[w][2];
    }
    if (fs | fx) {
      u64* fp = a.fp + (stripe % NSTRIPE) * 2;
      atomicAdd((unsigned long long*)&fp[0], (unsigned long long)fs);
      atomicXor((unsigned long long*)&fp[1], (unsigned long long)fx);
    }
    if (tt) {
      u64* sp = a.stat + (stripe % NSTRIPE) * NSTAT;
      atomicAdd((unsigned long long*)&sp[1], (unsigned long long)tt);
      atomicAdd((unsigned long long*)&sp[2], (unsigned long long)te);
      atomicAdd((unsigned long long*)&sp[5], (unsigned long long)tb);
    }
    u64 wb = 0;
    if (btot) {
      wb = atomicAdd((unsigned long long*)a.res_words, (unsigned long long)btot);
      if (wb + btot > a.res_cap) atomicOr(a.flags, FLAG_RES_OVERFLOW);
    }
    s_wbase = wb;
  }
  __syncthreads();
  const u64 wb = s_wbase;
  if (!btot || wb + btot > a.res_cap) return want;  // the same in every thread
  u64 boff = off;
#pragma unroll
  for (int w = 0; w < (int)CHECK_WAVES; ++w) boff += w < wv ? s_w[w] : 0ull;
  if (btot <= STAGE_WORDS) {
    if (want) write_record<K>(s_stage + boff, a, f, ch, nch);
    __syncthreads();
    for (u32 i = threadIdx.x; i < (u32)btot; i += CHECK_THREADS) a.res[wb + i] = s_stage[i];
  } else if (want) {
    write_record<K>(a.res + wb + boff, a, f, ch, nch);
  }
  return want;
}

// Checks rank r (one lane): unrank its combination into f and ch, test
// maximality.  Returns whether it is maximal (the caller emits it).  A rank
// still open after its first a.budget candidates goes to the deferred list
// (thin_check_deferred tests the rest, a wave per rank).
template <int K>
__device__ __forceinline__ bool check_rank(const ThinArgs& a, u64 r, const thin::Family*& f, u32 (&ch)[2 * K * K],
                                           u32& nch, thin::Cost& cost, u64& n_tested) {
  constexpr int NM = 2 * K * K;
  if (r >= a.r1) return false;
  u32 ctsz[NM];
  f = unrank_rank<K>(a, r, ch, ctsz, nch, cost);
  if (!f) return false;
  const int m = thin::side_a_maximal<NM>(a.g, fam_view(a, *f), ch, ctsz, nch, a.k, cost, a.budget);
  if (m < 0) {
    const u32 i = atomicAdd(a.defer_n, 1u);
    a.defer[i] = r;  // the list holds a launch's ranks
    return false;
  }
  ++n_tested;
  return m == 1;
}

// Deferred ranks, one wave each: every lane unranks the combination, the lanes
// split its candidates after the first a.budget (which the first pass tested),
// and lane 0 emits it when none can join.
// A fixed grid of waves takes deferred ranks from a device counter.
template <int K>
__global__ __launch_bounds__(256) void thin_check_deferred(ThinArgs a) {
  constexpr int NM = 2 * K * K;
  const int lane = threadIdx.x & 63;
  u64 n_emit = 0, n_tested = 0;
  thin::Cost cost;
  const u32 nd = __hip_atomic_load(a.defer_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (nd == 0) return;
  for (;;) {
    u32 i = 0;
    if (lane == 0) i = atomicAdd(a.dwork, 1u);
    i = (u32)wave_bcast_u64(i, 0);
    if (i >= nd) break;
    if (limit_reached(a)) break;
    const u64 t0 = a.item_log ? (u64)wall_clock64() : 0;
    const u64 r = a.defer[i];
    u32 ch[NM], ctsz[NM];
    u32 nch = 0;
    const thin::Family* f = unrank_rank<K>(a, r, ch, ctsz, nch, cost);
    if (!f) continue;  // the same rank in every lane
    thin::CandSegs cs;
    const thin::FamRef fr = fam_view(a, *f);
    thin::candidate_segments<NM>(a.g, fr, ch, ctsz, nch, a.k, cs, cost);
    u64 total = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) total += (u32)q < cs.n ? cs.len[q] : 0u;
    bool joins = false;
    for (u64 base = a.budget; base < total; base += 64) {
      const u64 idx = base + lane;
      if (idx < total) {
        // list and position of flattened candidate idx (constant indices into cs)
        u64 o = idx;
        const u32* lst = cs.lst[0];
        u32 known = cs.known[0];
        bool found = false;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (!found && (u32)q < cs.n) {
            if (o < cs.len[q]) { lst = cs.lst[q]; known = cs.known[q]; found = true; }
            else o -= cs.len[q];
          }
        }
        cost.bytes += 4;
        joins = thin::candidate_joins<NM>(a.g, fr, ch, ctsz, nch, a.k, cs, known, lst[o], cost);
      }
      if (__ballot(joins)) break;
    }
    const bool maximal = cs.decided >= 0 ? cs.decided == 1 : !__ballot(joins);
    if (lane == 0) ++n_tested;
    const bool em = emit_wave<K>(a, lane == 0 && maximal, f, ch, nch, blockIdx.x * 4 + (threadIdx.x >> 6), cost);
    if (em) ++n_emit;
    if (lane == 0) log_item(a, f, em, t0);
  }
  flush_check_stats<K>(a, cost, n_emit, n_tested);
}

// One lane per rank over the whole grid; a workgroup emits its results
// together (emit_block), so every thread reaches that call.
template <int K>
__global__ __launch_bounds__(CHECK_THREADS) void thin_check(ThinArgs a) {
  constexpr int NM = 2 * K * K;
  const u64 r = a.r0 + (u64)blockIdx.x * blockDim.x + threadIdx.x;
  u64 n_tested = 0;
  thin::Cost cost;
  const thin::Family* f = nullptr;
  u32 ch[NM], nch = 0;
  bool want = false;
  const u64 t0 = a.item_log ? (u64)wall_clock64() : 0;
  if (!limit_reached(a)) want = check_rank<K>(a, r, f, ch, nch, cost, n_tested);
  const bool em = emit_block<K>(a, want, f, ch, nch, n_tested != 0, cost);
  log_item(a, f, em, t0);
}

// Chunked form: a fixed grid of waves takes chunks of 64 consecutive ranks from
// a device counter until the launch's range is done, so a wave held by an
// expensive rank does not hold back the chunks after it.
template <int K>
__global__ __launch_bounds__(256) void thin_check_chunked(ThinArgs a) {
  constexpr int NM = 2 * K * K;
  const int lane = threadIdx.x & 63;
  u64 n_emit = 0, n_tested = 0;
  thin::Cost cost;
  for (;;) {
    u32 chunk = 0;
    if (lane == 0) chunk = atomicAdd(a.work, 1u);
    chunk = (u32)wave_bcast_u64(chunk, 0);
    const u64 c0 = a.r0 + (u64)chunk * 64;
    if (c0 >= a.r1) break;
    if (limit_reached(a)) break;
    const u64 t0 = a.item_log ? (u64)wall_clock64() : 0;
    const thin::Family* f = nullptr;
    u32 ch[NM], nch = 0;
    const bool want = check_rank<K>(a, c0 + lane, f, ch, nch, cost, n_tested);
    const bool em = emit_wave<K>(a, want, f, ch, nch, blockIdx.x * 4 + (threadIdx.x >> 6), cost);
    if (em) ++n_emit;
    log_item(a, f, em, t0);
  }
  flush_check_stats<K>(a, cost, n_emit, n_tested);
}

hipError_t launch_thin_build(const ThinArgs& a, hipStream_t s) {
  if (a.ntuples == 0) return hipSuccess;
  if (a.j >= 2 && a.j > a.k) {
    const dim3 g((a.ntuples + 3) / 4), b(256);
    switch (a.j) {
      case 2: hipLaunchKernelGGL(thin_build_wave<2>, g, b, 0, s, a); break;
      case 3: hipLaunchKernelGGL(thin_build_wave<3>, g, b, 0, s, a); break;
      case 4: hipLaunchKernelGGL(thin_build_wave<4>, g, b, 0, s, a); break;
      case 5: hipLaunchKernelGGL(thin_build_wave<5>, g, b, 0, s, a); break;
      default: hipLaunchKernelGGL(thin_build_wave<6>, g, b, 0, s, a); break;
    }
  } else if (a.j == 1 && a.ntuples < a.wave1_max) {
    const dim3 g((a.ntuples + 3) / 4), b(256);
    hipLaunchKernelGGL(thin_build_wave1, g, b, 0, s, a);
    hipLaunchKernelGGL(thin_scan_counts, dim3(1), dim3(1024), 0, s, a.tcount, a.ntuples, a.fam_n, a.fam_cap,
                       a.tflags);
    hipLaunchKernelGGL(thin_emit_wave1, dim3((a.ntuples + 255) / 256), b, 0, s, a);
  } else {
    const dim3 g((a.ntuples + 255) / 256), b(256);
    switch (a.j) {
      case 1: hipLaunchKernelGGL(thin_build_families<1>, g, b, 0, s, a); break;
      case 2: hipLaunchKernelGGL(thin_build_families<2>, g, b, 0, s, a); break;
      case 3: hipLaunchKernelGGL(thin_build_families<3>, g, b, 0, s, a); break;
      default: return hipErrorInvalidValue;  // j <= k <= 3 here
    }
  }
  return hipGetLastError();
}

// Guard tables (thin.h): one wave per family of the launch.  The lanes split
// the candidate lists of the plan, count each candidate's misses on the free
// list and insert the qualifying ones by compare-and-swap into a table taken
// from the arena; lane 0 then sets the family's guard mode (or drops a
// dominated family by setting its total to 0).  A family whose plan is over
// the size bounds, whose scan passes the work bound or whose table finds the
// arena full keeps the plain check.
struct DevGuardIns {
  u32* t;
  u32 lg;
  __device__ void operator()(u32 w, u32 m) {
    const u32 mask = (1u << lg) - 1u, word = w | (m << 30);
    for (u32 s = thin::nh_slot(w, lg);; s = (s + 1) & mask) {
      const u32 old = atomicCAS(&t[s], thin::GUARD_EMPTY, word);
      if (old == thin::GUARD_EMPTY || (old & thin::GUARD_VMASK) == w) return;
    }
  }
};

__device__ void guard_family(const ThinArgs& a, thin::Family* f, thin::Cost& cost) {
  const int lane = threadIdx.x & 63;
  if (f->total == 0) return;
  thin::GuardPlan p;
  thin::guard_plan(a.g, *f, a.k, a.guard_min_ranks, p, cost);
  if (!p.ok) return;
  const u32 lg = thin::guard_log2(p.ctot);
  u64 at = 0;
  if (p.ctot) {
    if (lane == 0) {
      DevArena ar{a.arena, a.arena_cap, a.arena_used};
      at = ar.alloc(1ull << lg);
    }
    at = wave_bcast_u64(at, 0);
  }
  if (at == thin::SAT) return;
  u32* t = p.ctot ? a.arena + at : nullptr;
  for (u32 s = lane; p.ctot && s < (1u << lg); s += 64) t[s] = thin::GUARD_EMPTY;
  __threadfence();
  DevGuardIns ins{t, lg};
  bool over = false;
  const u32 gm = thin::guard_scan(a.g, *f, a.k, p, (u32)lane, 64u, ins, over, cost);
  const u32 gmin_t = (u32)wave_min_u64((u64)gm);
  if (__ballot(over) != 0) return;
  if (lane == 0) {
    bool dominated = false;
    u32 gmin = 0;
    const u32 mode = thin::guard_finish(a.g, f->j, f->nfree, a.k, p, gmin_t, dominated, gmin);
    if (dominated) f->total = 0;
    f->gmin = gmin;
    f->s1cap = thin::guard_s1cap(f->nfree, a.k, p);
    f->glog = lg;
    f->gtab = (u64)t;
    f->gmode = mode;
  }
}

// A fixed grid of waves takes the families of the launch (their number is read
// on the device, so the guard build follows the construction kernels with no
// host round trip).
__global__ __launch_bounds__(256) void thin_guard(ThinArgs a) {
  const int lane = threadIdx.x & 63;
  const u32 wv = blockIdx.x * 4 + (threadIdx.x >> 6), nwv = gridDim.x * 4;
  u32 nf = __hip_atomic_load(a.fam_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (nf > a.fam_cap) nf = a.fam_cap;
  thin::Cost cost;
  for (u32 i = wv; i < nf; i += nwv) guard_family(a, a.fams + i, cost);
  const u64 b = wave_sum_u64(cost.bytes);
  if (lane == 0 && b) atomicAdd((unsigned long long*)&a.stat[(wv % NSTRIPE) * NSTAT + 4], (unsigned long long)b);
}

hipError_t launch_thin_guard(const ThinArgs& a, hipStream_t s, u32 blocks) {
  hipLaunchKernelGGL(thin_guard, dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

// One thread per index entry: binary search of the family prefix sums.
__global__ __launch_bounds__(256) void thin_rank_index(const u64* prefix, u32 nfam, u64 r0, u64 r1, u32 nb,
                                                       u32* ridx, u32* zero3) {
  const u32 b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < 3) zero3[b] = 0;
  if (b > nb) return;
  const u64 rb = r0 + ((u64)b << RIDX_LOG);
  const u64 r = rb < r1 ? rb : r1 - 1;
  u32 lo = 0, hi = nfam;
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (prefix[mid] <= r) lo = mid; else hi = mid;
  }
  ridx[b] = lo;
}

hipError_t launch_thin_rank_index(const u64* prefix, u32 nfam, u64 r0, u64 r1, u32* ridx, u32* zero3,
                                  hipStream_t s) {
  const u32 nb = (u32)((r1 - r0 + (1ull << RIDX_LOG) - 1) >> RIDX_LOG);
  hipLaunchKernelGGL(thin_rank_index, dim3((nb + 1 + 255) / 256), dim3(256), 0, s, prefix, nfam, r0, r1, nb, ridx,
                     zero3);
  return hipGetLastError();
}

hipError_t launch_thin_scan(const thin::Family* fams, u32 nfam, u64* prefix, u64 base, u64 cap, hipStream_t s) {
  hipLaunchKernelGGL(thin_scan_totals, dim3(1), dim3(thin::TSCAN), 0, s, fams, nfam, (const u32*)nullptr, 0u, prefix,
                     base, cap, (u64*)nullptr);
  return hipGetLastError();
}

hipError_t launch_thin_scan_dev(const thin::Family* fams, const u32* nfam_p, u32 fam_cap, u64* prefix, u64 base,
                                u64 cap, u64* total_out, hipStream_t s) {
  hipLaunchKernelGGL(thin_scan_totals, dim3(1), dim3(thin::TSCAN), 0, s, fams, 0u, nfam_p, fam_cap, prefix, base, cap,
                     total_out);
  return hipGetLastError();
}

// One wave per work item (a hashed vertex and a chunk of NH_CHUNK of its
// neighbours, so hubs spread over many waves): lanes insert the neighbours by
// compare-and-swap with linear probing (neighbours are distinct, the set has
// >= 2d slots).
__global__ __launch_bounds__(256) void build_neighbour_hash(const u32* off, const u32* adj, const u32* items, u32 nv,
                                                            const u64* nh_off, u32* nh) {
  const u32 wv = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (wv >= nv) return;
  const u32 v = items[2 * wv], c0 = items[2 * wv + 1];
  const u32 o = off[v], d = off[v + 1] - o;
  const u32 lg = thin::nh_log2(d), mask = (1u << lg) - 1u;
  u32* t = nh + nh_off[v];
  const u32 c1 = d - c0 < NH_CHUNK ? d : c0 + NH_CHUNK;
  for (u32 i = c0 + lane; i < c1; i += 64) {
    const u32 x = adj[o + i];
    u32 sl = thin::nh_slot(x, lg);
    while (atomicCAS(&t[sl], thin::NH_EMPTY, x) != thin::NH_EMPTY) sl = (sl + 1) & mask;
  }
}

hipError_t launch_build_neighbour_hash(const u32* off, const u32* adj, const u32* items, u32 nv, const u64* nh_off,
                                       u32* nh, hipStream_t s) {
  if (nv == 0) return hipSuccess;
  hipLaunchKernelGGL(build_neighbour_hash, dim3((nv + 3) / 4), dim3(256), 0, s, off, adj, items, nv, nh_off, nh);
  return hipGetLastError();
}

hipError_t launch_thin_check_deferred(const ThinArgs& a, hipStream_t s, u32 blocks) {
  const dim3 g(blocks), b(256);
  switch (a.k) {
    case 1: hipLaunchKernelGGL(thin_check_deferred<1>, g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL(thin_check_deferred<2>, g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL(thin_check_deferred<3>, g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_thin_check(const ThinArgs& a, hipStream_t s, u32 max_blocks, u32 form) {
  if (a.r1 <= a.r0) return hipSuccess;
  const u64 n = a.r1 - a.r0;
  const u64 need = (n + 255) / 256;
  const dim3 b(256);
  if (form == CHECK_CHUNKED) {
    const dim3 g((u32)(need < max_blocks ? need : max_blocks));
    switch (a.k) {
      case 1: hipLaunchKernelGGL(thin_check_chunked<1>, g, b, 0, s, a); break;
      case 2: hipLaunchKernelGGL(thin_check_chunked<2>, g, b, 0, s, a); break;
      case 3: hipLaunchKernelGGL(thin_check_chunked<3>, g, b, 0, s, a); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    const dim3 g((u32)((n + CHECK_THREADS - 1) / CHECK_THREADS)), bc(CHECK_THREADS);
    switch (a.k) {
      case 1: hipLaunchKernelGGL(thin_check<1>, g, bc, 0, s, a); break;
      case 2: hipLaunchKernelGGL(thin_check<2>, g, bc, 0, s, a); break;
      case 3: hipLaunchKernelGGL(thin_check<3>, g, bc, 0, s, a); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

}  // namespace gmk
