// Thin-result kernels (see thin.h): family construction (one lane per tuple)
// and concrete-combination checking with emission (one lane per rank).
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

__global__ __launch_bounds__(256) void thin_build_families(ThinArgs a) {
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntuples) return;
  DevArena ar{a.arena, a.arena_cap, a.arena_used};
  DevFamilies fo{a.fams, a.fam_cap, a.fam_n, a.tflags};
  thin::Cost cost;
  if (!thin::build_families(a.g, a.tuples + (u64)t * a.j, a.j, a.k, a.min_b, t, ar, fo, cost))
    atomicOr(a.tflags, FLAG_RES_OVERFLOW);
  atomicAdd((unsigned long long*)&a.stat[(t / 64 % NSTRIPE) * NSTAT + 4], (unsigned long long)cost.bytes);
}

__global__ __launch_bounds__(256) void thin_check(ThinArgs a) {
  const u64 r = a.r0 + (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool emit = false;
  u32 ch[thin::MAXJ * 3], ctsz[thin::MAXJ * 3];
  u32 nch = 0;
  const thin::Family* f = nullptr;
  thin::Cost cost;
  if (r < a.r1 && !(__hip_atomic_load(a.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FLAG_LIMIT)) {
    // family containing rank r
    u32 lo = 0, hi = a.nfam;
    while (hi - lo > 1) {
      const u32 mid = (lo + hi) >> 1;
      if (a.prefix[mid] <= r) lo = mid; else hi = mid;
    }
    f = a.fams + lo;
    u64 rr = r - a.prefix[lo];
    for (u32 t = 0; t < f->ntypes; ++t) {
      const u64 radix = thin::binom(f->psize[t], f->cnt[t]);
      const u64 rt = rr % radix;
      rr /= radix;
      u32 idx[3];
      thin::unrank(rt, f->cnt[t], f->psize[t], idx);
      const u32 tsz = thin::popc(f->mask[t]);
      for (u32 q = 0; q < f->cnt[t]; ++q) {
        const u32 y = f->poff[t] == thin::FAR
                          ? thin::far_select(a.arena + f->nl_off, f->nl_size, a.g.b_lo, idx[q], cost)
                          : a.arena[f->poff[t] + idx[q]];
        cost.bytes += 4;
        ch[nch] = y; ctsz[nch] = tsz; ++nch;
      }
    }
    cost.bytes += 8ull * 8 + 16ull * f->ntypes;  // prefix probes, family header and type fields
    emit = thin::side_a_maximal(a.g, f->L, f->j, a.arena + f->free_off, f->nfree, ch, ctsz, nch, a.k, cost);
  }
  // per-lane emission (records are independent)
  u64 h = 0;
  if (emit) {
    const u64 slot = atomicAdd((unsigned long long*)a.res_count, 1ull);
    if (slot >= a.limit) {
      atomicOr(a.flags, FLAG_LIMIT);
      emit = false;
    } else {
      if (slot + 1 == a.limit) atomicOr(a.flags, FLAG_LIMIT);
      // sort the chosen vertices (few) for the merge with the free list
      for (u32 i = 1; i < nch; ++i)
        for (u32 q = i; q > 0 && ch[q - 1] > ch[q]; --q) { const u32 tmp = ch[q]; ch[q] = ch[q - 1]; ch[q - 1] = tmp; }
      const u32 nr = f->nfree + nch;
      const u32* fr = a.arena + f->free_off;
      for (u32 i = 0; i < f->j; ++i) h += mix64(f->L[i]);
      for (u32 i = 0; i < f->nfree; ++i) h += mix64(fr[i]);
      for (u32 i = 0; i < nch; ++i) h += mix64(ch[i]);
      const u32 nl = f->j;
      const u64 fpv = fmix64(h + (u64)(a.swap_sides ? nr : nl) * FP_A + (u64)(a.swap_sides ? nl : nr) * FP_B);
      u64* fs = a.fp + ((blockIdx.x * 4 + (threadIdx.x >> 6)) % NSTRIPE) * 2;
      atomicAdd((unsigned long long*)&fs[0], (unsigned long long)fpv);
      atomicXor((unsigned long long*)&fs[1], (unsigned long long)fpv);
      cost.bytes += 4ull * f->nfree;  // free list read for the fingerprint / record
      if (a.keep) {
        const u64 len = 2ull + nl + nr;
        cost.bytes += 4ull * len;
        const u64 pos = atomicAdd((unsigned long long*)a.res_words, (unsigned long long)len);
        if (pos + len > a.res_cap) {
          atomicOr(a.flags, FLAG_RES_OVERFLOW);
        } else {
          // tuple side and generated side, left side first
          u64 at_t = a.swap_sides ? pos + 1 + nr + 1 : pos + 1;
          u64 at_r = a.swap_sides ? pos + 1 : pos + 1 + nl + 1;
          a.res[a.swap_sides ? pos + 1 + nr : pos] = nl;
          a.res[a.swap_sides ? pos : pos + 1 + nl] = nr;
          for (u32 i = 0; i < nl; ++i) a.res[at_t + i] = f->L[i];
          u32 p = 0, q = 0;
          for (u32 i = 0; i < nr; ++i) {
            u32 v;
            if (q >= nch || (p < f->nfree && fr[p] < ch[q])) v = fr[p++];
            else v = ch[q++];
            a.res[at_r + i] = v;
          }
        }
      }
    }
  }
  (void)lane;
  const u64 e = emit ? 1 : 0;
  const u64 tested = (r < a.r1 && f) ? 1 : 0;
  const u64 se = wave_sum_u64(e), stt = wave_sum_u64(tested), sb = wave_sum_u64(cost.bytes);
  if (lane == 0 && stt) {
    u64* sp = a.stat + ((blockIdx.x * 4 + (threadIdx.x >> 6)) % NSTRIPE) * NSTAT;
    atomicAdd((unsigned long long*)&sp[1], (unsigned long long)stt);
    atomicAdd((unsigned long long*)&sp[2], (unsigned long long)se);
    atomicAdd((unsigned long long*)&sp[5], (unsigned long long)sb);
  }
}

hipError_t launch_thin_build(const ThinArgs& a, hipStream_t s) {
  if (a.ntuples == 0) return hipSuccess;
  hipLaunchKernelGGL(thin_build_families, dim3((a.ntuples + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_thin_check(const ThinArgs& a, hipStream_t s) {
  if (a.r1 <= a.r0) return hipSuccess;
  const u64 n = a.r1 - a.r0;
  hipLaunchKernelGGL(thin_check, dim3((u32)((n + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace gmk
