// Seed partitions' vertex sets built on the device: the role of BPPivot's
// constructSets and pruneCX (baselines/MaximalBiPlex/src/biplex/biplex.cpp:
// 227-393), as the host restatement csrc/seeds.cpp builds them, one workgroup
// per seed.
//
// For a seed s (side us):  C_os = N(s);  the 2-hop vertices are C_us (after s
// in the degeneracy order) or X_us (before it);  then pruning;  then, when
// k >= 1, the 3-hop vertices of the other side with lb - k or more candidate
// neighbours join C_os;  then pruning again.  Pruning removes
//   a C_os vertex with fewer than lb - k neighbours among C_us and s,
//   a C_us or X_us vertex with fewer than lb - 2k neighbours in C_os,
// and the partition is dropped when s keeps fewer than lb - k neighbours in
// C_os.  Removal only lowers the counts, so the surviving sets are the
// greatest ones meeting every bound, whatever the order of removals: the
// workgroup removes in parallel rounds (C_os, then C_us and X_us, until a round
// removes nothing) and reaches exactly the sets of the host's sequential loop.
// In a round the counts tested are stable: C_os counts fall only in the C_us
// round, C_us / X_us counts only in the C_os round.
//
// Per workgroup, an HBM slice holds a mark and a count per vertex of the
// graph (zero between seeds: every vertex a seed touches is on one of its
// lists and is cleared at the end) and the lists.  The surviving sides are
// sorted (bitonic network over a power-of-two scratch padded with ~0) and
// written to an output arena: the seed's side first (flag in the top two
// bits: 0 C, 1 X, 2 the seed), then the other side.
#include "gmk_device.h"

namespace gmk {

enum : u32 { SM_NONE = 0, SM_CS = 1, SM_XS = 2, SM_CO_N = 3, SM_SEED = 4, SM_CO_3 = 5, SM_DEAD = 6, SM_SEEN3 = 7 };

__device__ __forceinline__ bool same_member(u32 m) { return m == SM_CS || m == SM_XS || m == SM_SEED; }
__device__ __forceinline__ bool co_member(u32 m) { return m == SM_CO_N || m == SM_CO_3; }

// Appends v to a workgroup list (count in LDS).
__device__ __forceinline__ void push(u32* list, u32* n, u32 v) { list[atomicAdd(n, 1u)] = v; }

// Bitonic sort of list[0, len) ascending, through buf (len rounded up to a
// power of two, padding ~0).
__device__ void wg_sort(u32* list, u32 len, u32* buf) {
  u32 P = 1;
  while (P < len) P <<= 1;
  for (u32 i = threadIdx.x; i < P; i += blockDim.x) buf[i] = i < len ? list[i] : 0xFFFFFFFFu;
  __syncthreads();
  for (u32 kk = 2; kk <= P; kk <<= 1)
    for (u32 j = kk >> 1; j > 0; j >>= 1) {
      for (u32 i = threadIdx.x; i < P; i += blockDim.x) {
        const u32 p = i ^ j;
        if (p > i) {
          const u32 a = buf[i], b = buf[p];
          if (((i & kk) == 0) ? (a > b) : (a < b)) { buf[i] = b; buf[p] = a; }
        }
      }
      __syncthreads();
    }
  for (u32 i = threadIdx.x; i < len; i += blockDim.x) list[i] = buf[i];
  __syncthreads();
}

__global__ __launch_bounds__(256) void seed_sets(SeedArgs a) {
  __shared__ u32 n_co, n_cs, n_xs, n_cand, changed, n_same, n_other, status;
  __shared__ unsigned long long at_s;
  const u32 tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  u32* mark = a.scratch + (u64)blockIdx.x * 8 * a.n_max;
  int* cnt = (int*)(mark + a.n_max);
  u32* CO = mark + 2 * (u64)a.n_max;
  u32* CS = mark + 3 * (u64)a.n_max;
  u32* XS = mark + 4 * (u64)a.n_max;
  u32* CAND = mark + 5 * (u64)a.n_max;
  u32* SORT = mark + 6 * (u64)a.n_max;  // 2 n_max words
  u64 bytes = 0;
  for (u32 t = blockIdx.x; t < a.n_tasks; t += gridDim.x) {
    const SeedTask T = a.tasks[t];
    const u32* off = a.off + a.off_base[T.graph];
    const u32* adj = a.adj + a.adj_base[T.graph];
    const u32* order = a.order + a.off_base[T.graph];
    const u32 nl_g = a.n_left[T.graph];
    const u32 s = T.seed;
    const u32 us = s < nl_g ? 0u : 1u;
    const int need_os = (int)T.lb - (int)T.k, need_us = (int)T.lb - 2 * (int)T.k;
    if (tid == 0) { n_co = n_cs = n_xs = n_cand = 0; status = 1; }
    __syncthreads();
    // C_os = N(s)
    const u32 os0 = off[s], ds = off[s + 1] - os0;
    if (tid == 0) mark[s] = SM_SEED;
    for (u32 i = tid; i < ds; i += blockDim.x) {
      const u32 v = adj[os0 + i];
      mark[v] = SM_CO_N;
      CO[i] = v;
    }
    if (tid == 0) n_co = ds;
    bytes += 4ull * ds;
    __syncthreads();
    // 2-hop: C_us (after s in the order) and X_us (before it), each vertex once
    const u32 ord_s = order[s];
    for (u32 i = wv; i < ds; i += 4) {
      const u32 v = CO[i];
      const u32 ov = off[v], dv = off[v + 1] - ov;
      for (u32 j = lane; j < dv; j += 64) {
        const u32 w = adj[ov + j];
        const u32 m = order[w] > ord_s ? SM_CS : SM_XS;
        if (atomicCAS(&mark[w], (u32)SM_NONE, m) == SM_NONE) push(m == SM_CS ? CS : XS, m == SM_CS ? &n_cs : &n_xs, w);
      }
      bytes += 12ull * dv;
    }
    __syncthreads();
    // counts: same-side members' C_os neighbours; C_os vertices' C_us neighbours
    for (u32 i = wv; i < n_co; i += 4) {
      const u32 y = CO[i];
      const u32 oy = off[y], dy = off[y + 1] - oy;
      u32 c = 0;
      for (u32 j = lane; j < dy; j += 64) {
        const u32 x = adj[oy + j];
        const u32 m = mark[x];
        if (same_member(m)) atomicAdd(&cnt[x], 1);
        c += m == SM_CS ? 1u : 0u;
      }
      c = (u32)wave_sum_u64(c);
      if (lane == 0 && c) atomicAdd(&cnt[y], (int)c);
      bytes += 12ull * dy;
    }
    __syncthreads();
    for (int phase = 0; phase < 2; ++phase) {
      if (phase == 1) {
        // 3-hop vertices of the other side (k >= 1): candidates adjacent to
        // surviving C_us vertices, counted once per such neighbour
        if (T.k == 0 || cnt[s] < need_os) break;
        for (u32 i = wv; i < n_cs; i += 4) {
          const u32 w = CS[i];
          if (mark[w] != SM_CS) continue;
          const u32 ow = off[w], dw = off[w + 1] - ow;
          for (u32 j = lane; j < dw; j += 64) {
            const u32 y = adj[ow + j];
            if (atomicCAS(&mark[y], (u32)SM_NONE, (u32)SM_SEEN3) == SM_NONE) push(CAND, &n_cand, y);
            if (mark[y] == SM_SEEN3) atomicAdd(&cnt[y], 1);
          }
          bytes += 12ull * dw;
        }
        __syncthreads();
        const u32 nc = n_cand;
        for (u32 i = wv; i < nc; i += 4) {
          const u32 y = CAND[i];
          const bool join = cnt[y] + (int)T.k >= (int)T.lb;
          if (lane == 0) {
            mark[y] = join ? SM_CO_3 : SM_DEAD;
            if (join) push(CO, &n_co, y);
          }
          if (join) {
            const u32 oy = off[y], dy = off[y + 1] - oy;
            for (u32 j = lane; j < dy; j += 64) {
              const u32 x = adj[oy + j];
              if (same_member(mark[x])) atomicAdd(&cnt[x], 1);
            }
            bytes += 8ull * dy;
          }
        }
        __syncthreads();
      }
      // pruning to the greatest sets meeting the bounds
      for (;;) {
        if (tid == 0) changed = 0;
        __syncthreads();
        for (u32 i = wv; i < n_co; i += 4) {
          const u32 y = CO[i];
          int die = 0;
          if (lane == 0) {
            const u32 m = mark[y];
            die = co_member(m) && cnt[y] + (m == SM_CO_N ? 1 : 0) < need_os;
            if (die) { mark[y] = SM_DEAD; changed = 1; }
          }
          die = __shfl(die, 0, 64);
          if (die) {
            const u32 oy = off[y], dy = off[y + 1] - oy;
            for (u32 j = lane; j < dy; j += 64) {
              const u32 x = adj[oy + j];
              if (same_member(mark[x])) atomicSub(&cnt[x], 1);
            }
            bytes += 8ull * dy;
          }
        }
        __syncthreads();
        for (u32 i = wv; i < n_cs; i += 4) {
          const u32 w = CS[i];
          int die = 0;
          if (lane == 0) {
            die = mark[w] == SM_CS && cnt[w] < need_us;
            if (die) { mark[w] = SM_DEAD; changed = 1; }
          }
          die = __shfl(die, 0, 64);
          if (die) {
            const u32 ow = off[w], dw = off[w + 1] - ow;
            for (u32 j = lane; j < dw; j += 64) {
              const u32 y = adj[ow + j];
              if (co_member(mark[y])) atomicSub(&cnt[y], 1);
            }
            bytes += 8ull * dw;
          }
        }
        for (u32 i = tid; i < n_xs; i += blockDim.x) {
          const u32 x = XS[i];
          if (mark[x] == SM_XS && cnt[x] < need_us) { mark[x] = SM_DEAD; changed = 1; }
        }
        __syncthreads();
        const u32 more = changed;
        __syncthreads();  // every lane has read `changed` before it is cleared
        if (!more) break;
      }
    }
    // output: the surviving sides, sorted, in the arena
    const bool alive = cnt[s] >= need_os;
    if (tid == 0) { n_same = 0; n_other = 0; }
    __syncthreads();
    if (alive) {
      u32 cs = 0, co = 0;
      for (u32 i = tid; i < n_cs; i += blockDim.x) cs += mark[CS[i]] == SM_CS ? 1u : 0u;
      for (u32 i = tid; i < n_xs; i += blockDim.x) cs += mark[XS[i]] == SM_XS ? 1u : 0u;
      for (u32 i = tid; i < n_co; i += blockDim.x) co += co_member(mark[CO[i]]) ? 1u : 0u;
      cs = (u32)wave_sum_u64(cs);
      co = (u32)wave_sum_u64(co);
      if (lane == 0) { atomicAdd(&n_same, cs); atomicAdd(&n_other, co); }
    }
    __syncthreads();
    const u32 ns = alive ? n_same + 1 : 0, no = n_other;
    if (tid == 0) {
      status = alive ? 1u : 0u;
      at_s = 0;
      if (alive && (ns > a.side_cap || no > a.side_cap)) status = 3;
      if (status == 1) {
        const unsigned long long at = atomicAdd(a.arena_used, (unsigned long long)(ns + no));
        if (at + ns + no > a.arena_cap) status = 2;
        at_s = at;
      }
    }
    __syncthreads();
    if (status == 1) {
      u32* same = a.arena + at_s;
      u32* other = same + ns;
      if (tid == 0) { n_same = 1; n_other = 0; same[0] = s; }
      __syncthreads();
      for (u32 i = tid; i < n_cs; i += blockDim.x)
        if (mark[CS[i]] == SM_CS) same[atomicAdd(&n_same, 1u)] = CS[i];
      for (u32 i = tid; i < n_xs; i += blockDim.x)
        if (mark[XS[i]] == SM_XS) same[atomicAdd(&n_same, 1u)] = XS[i];
      for (u32 i = tid; i < n_co; i += blockDim.x)
        if (co_member(mark[CO[i]])) other[atomicAdd(&n_other, 1u)] = CO[i];
      __syncthreads();
      wg_sort(same, ns, SORT);
      wg_sort(other, no, SORT);
      for (u32 i = tid; i < ns; i += blockDim.x) {
        const u32 v = same[i], m = mark[v];
        same[i] = v | ((m == SM_SEED ? 2u : (m == SM_XS ? 1u : 0u)) << 30);
      }
      bytes += 8ull * (ns + no) * 4;
    }
    if (tid == 0) {
      SeedOut o;
      o.status = status;
      o.us = us;
      o.nl = us ? no : ns;
      o.nr = us ? ns : no;
      o.at = at_s;
      a.out[t] = o;
    }
    // clear every vertex this seed touched
    __syncthreads();
    if (tid == 0) { mark[s] = 0; cnt[s] = 0; }
    for (u32 i = tid; i < n_co; i += blockDim.x) { mark[CO[i]] = 0; cnt[CO[i]] = 0; }
    for (u32 i = tid; i < n_cs; i += blockDim.x) { mark[CS[i]] = 0; cnt[CS[i]] = 0; }
    for (u32 i = tid; i < n_xs; i += blockDim.x) { mark[XS[i]] = 0; cnt[XS[i]] = 0; }
    for (u32 i = tid; i < n_cand; i += blockDim.x) { mark[CAND[i]] = 0; cnt[CAND[i]] = 0; }
    __syncthreads();
  }
  const u64 b = wave_sum_u64(bytes);
  if (lane == 0 && b) atomicAdd(a.bytes, (unsigned long long)b);
}

hipError_t launch_seed_sets(const SeedArgs& a, u32 grid, hipStream_t s) {
  if (a.n_tasks == 0) return hipSuccess;
  hipLaunchKernelGGL(seed_sets, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace gmk
