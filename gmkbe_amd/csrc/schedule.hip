// Learned schedule of single-graph runs.  A work item is a one-vertex tuple of
// the thin enumerator (a vertex u of side A with its families: u with N(u) and
// up to k vertices it misses).  The item cost model (the FP32 MLP of mlp.hip,
// trained on the yield and device time the method logs per item, see
// gmk_probe_items) predicts log(results per ms of device lane time) from the
// item's features; a limited run takes the items with the best prediction
// first.  The model only reorders work: every item is still enumerated
// exactly, and a run without a model keeps the degree order of the IDs.
//
// Device work, on the run's stream with one synchronisation at the end: one
// lane per item computes its features and the forward pass (item_score), every
// 64th score is sampled and a radix select over the samples sets the threshold
// of about 1.5 `want` items (select_threshold), the items above it are
// gathered (score_select) and sorted best first by rank (sort_front).  The
// host receives the sorted front.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gmk_host.h"

using namespace gmk;

namespace {

constexpr int IN = GMK_MLP_IN, H = GMK_MLP_HID;
constexpr u32 LOG2_SAMPLE = 4;
constexpr u32 SAMPLE = 1u << LOG2_SAMPLE;  // neighbours sampled per item for the degree moments
constexpr u32 SSTEP_MIN = 64;              // every sstep-th item's key is sampled for the selection threshold
constexpr u32 SEL_PER_LANE = 32;           // samples per lane of select_threshold (at most 32768 samples)
constexpr u32 FRONT_CAP = 8192;            // gathered items, at most (sorted in LDS: 64 KiB)
constexpr u32 SELECT_ITEMS = 4096;         // items per workgroup of score_select

struct ItemArgs {
  const u32* off;
  const u32* adj;
  const u32* deg;   // vertex degrees
  u32 first, n;     // items: vertices first .. first + n - 1
  u32 nb;           // vertices on side B
  float k;
  const float* P;   // model parameters (gmk::mlp_pack order), then mu[IN], sd[IN]; null: features only
  float* feat;      // n x IN features, or null
  u32* key;         // order-preserving score keys, or null
  u32* sample;      // key of every sstep-th item, or null
  float* score;     // predictions, or null
  u32 sstep;        // sample stride, a power of two (with sample)
};

// Order-preserving map float -> u32 (larger float, larger key; NaN lowest).
__device__ __forceinline__ u32 score_key(float y) {
  if (y != y) return 0u;
  const u32 b = __float_as_uint(y);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Features of item u (gmk_item_features documents them).  Degree moments over
// at most SAMPLE neighbours spread evenly over N(u); the two-hop count is the
// sampled mean scaled to d(u).
__device__ __forceinline__ void item_features(const ItemArgs& a, u32 u, float (&x)[IN]) {
  const u32 o = a.off[u], d = a.off[u + 1] - o;
  const u32 ns = d < SAMPLE ? d : SAMPLE;
  // the sampled neighbours' degrees, every load of a round issued together
  u32 nb[SAMPLE], dv[SAMPLE];
#pragma unroll
  for (u32 i = 0; i < SAMPLE; ++i) {
    // i * d / ns with ns == SAMPLE a power of two when d > SAMPLE
    const u32 e = ns == d ? i : (u32)(((u64)i * d) >> LOG2_SAMPLE);
    nb[i] = i < ns ? a.adj[o + e] : 0u;
  }
#pragma unroll
  for (u32 i = 0; i < SAMPLE; ++i) dv[i] = i < ns ? a.deg[nb[i]] : 0u;
  // mean and variance by Welford's update (no cancellation when the degrees agree)
  float mean = 0.f, m2 = 0.f, mx = 0.f;
  u32 mn = 0xFFFFFFFFu;
  double hop = 0;
#pragma unroll
  for (u32 i = 0; i < SAMPLE; ++i) {
    if (i < ns) {
      const float l = __logf(1.f + (float)dv[i]);
      const float delta = l - mean;
      mean += __fdividef(delta, (float)(i + 1));
      m2 += delta * (l - mean);
      mx = l > mx ? l : mx;
      mn = dv[i] < mn ? dv[i] : mn;
      hop += (double)(dv[i] - 1);
    }
  }
  const float var = ns ? m2 / (float)ns : 0.f;
  x[0] = log1pf((float)d);
  x[1] = log1pf((float)(a.nb - d));
  x[2] = mean;
  x[3] = mx;
  x[4] = var > 0.f ? sqrtf(var) : 0.f;
  x[5] = log1pf(ns ? (float)(hop * (double)d / (double)ns) : 0.f);
  x[6] = log1pf(ns ? (float)mn : 0.f);
  x[7] = a.k;
}

// x split into bf16 parts, x ~ hi + lo (each rounded to nearest even by the
// hardware conversion), as bits.
__device__ __forceinline__ void split_bf16(float x, short& hi, short& lo) {
  const __bf16 h = (__bf16)x;
  hi = __builtin_bit_cast(short, h);
  lo = __builtin_bit_cast(short, (__bf16)(x - (float)h));
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// v_mfma_f32_32x32x16_bf16: D += A B with A 32 x 16, B 16 x 32, the operands
// given as bf16 bits
__device__ __forceinline__ f32x16 mfma16(s16x8 a, s16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

// One lane per item: features, then (with a model) the forward pass, 64 items
// per wave, both weight layers on the matrix cores as Z^T = W^T X^T over
// tiles of 32 units x 32 items, each operand split into bf16 high and low
// parts (a b ~ ah bh + ah bl + al bh, three products per step, FP32
// accumulation: about 2^-15 relative error per product).  The first layer
// (8 inputs) uses v_mfma_f32_32x32x8_bf16, the second v_mfma_f32_32x32x16_bf16.
// Operand layout (lane l, K = 4 or 8 values per lane, j < K):
// A[i = l % 32][k = K (l / 32) + j], B[k = K (l / 32) + j][n = l % 32]; result
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32], r = 0..15, for both.  Lane l feeds
// item n = 32 t + l % 32 of tile t: its own item or that of lane l ^ 32, whose
// features it takes by a shuffle.  A first-layer result tile holds, in lane
// half h, units 32 m + 4 h + {0..3} + 8 {0..3}; after bias and ReLU, its
// registers 8 q .. 8 q + 7 are the B operand of a 16-unit second-layer step as
// they stand, once the A operand (W2^T, read from LDS) takes its columns in the
// same unit order (the instruction sums over the K pairs of A and B, whatever
// units they stand for).  The weights sit in LDS transposed, split into their
// two bf16 parts.  The output layer sums each item's 64 units over the lane
// pair (l, l ^ 32).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void item_score(ItemArgs a) {
  __shared__ float b1[H], b2[H], w3[H], mu[IN], sd[IN];
  __shared__ float b3;
  __shared__ __align__(16) short w1h[H][IN + 4], w1l[H][IN + 4];  // W1^T: [unit][input]
  __shared__ __align__(16) short w2h[H][H + 4], w2l[H][H + 4];    // W2^T: [output unit][input unit]
  if (a.P) {
    const float* P = a.P;
    constexpr int OW1 = 0, OB1 = OW1 + IN * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
    for (int i = threadIdx.x; i < IN * H; i += blockDim.x) {
      const float w = P[OW1 + i];  // w1[input i / H][unit i % H]
      split_bf16(w, w1h[i % H][i / H], w1l[i % H][i / H]);
    }
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) {
      const float w = P[OW2 + i];  // w2[input i / H][output i % H]
      split_bf16(w, w2h[i % H][i / H], w2l[i % H][i / H]);
    }
    for (int i = threadIdx.x; i < H; i += blockDim.x) { b1[i] = P[OB1 + i]; b2[i] = P[OB2 + i]; w3[i] = P[OW3 + i]; }
    if (threadIdx.x < IN) { mu[threadIdx.x] = P[OB3 + 1 + threadIdx.x]; sd[threadIdx.x] = P[OB3 + 1 + IN + threadIdx.x]; }
    if (threadIdx.x == 0) b3 = P[OB3];
    __syncthreads();
  }
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < a.n;
  float x[IN];
  if (live) item_features(a, a.first + i, x);
  if (live && a.feat) {
#pragma unroll
    for (int q = 0; q < IN; ++q) a.feat[(u64)i * IN + q] = x[q];
  }
  if (!a.P) return;
  // every lane of the wave takes part in the matrix products from here
#pragma unroll
  for (int q = 0; q < IN; ++q) x[q] = live ? (x[q] - mu[q]) / sd[q] : 0.f;
  const int lane = threadIdx.x & 63, half = lane >> 5, col = lane & 31;
  float xp[IN];  // features of lane l ^ 32's item
#pragma unroll
  for (int q = 0; q < IN; ++q) xp[q] = __shfl_xor(x[q], 32, 64);
  // first layer: B operand of tile t = features 4 half + j of item 32 t + l % 32
  bf16x4 xh[2], xl[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float v = t == half ? (half ? x[4 + j] : x[j]) : (half ? xp[4 + j] : xp[j]);
      short hi, lo;
      split_bf16(v, hi, lo);
      xh[t][j] = hi;
      xl[t][j] = lo;
    }
  // Output tiles of the second layer one at a time (32 units each), the first
  // layer recomputed per tile, 32 units at a time: at most two result tiles
  // per item tile are live, so four waves fit on a SIMD.
  float part[2] = {0.f, 0.f};
#pragma unroll
  for (int mo = 0; mo < 2; ++mo) {
    f32x16 acc[2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[t][r] = 0.f;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
      // first layer, units 32 mi .. 32 mi + 31
      f32x16 z1[2];
      const bf16x4 ah1 = *(const bf16x4*)&w1h[32 * mi + col][4 * half];
      const bf16x4 al1 = *(const bf16x4*)&w1l[32 * mi + col][4 * half];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        f32x16 c;
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] = 0.f;
        c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah1, xh[t], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah1, xl[t], c, 0, 0, 0);
        z1[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(al1, xh[t], c, 0, 0, 0);
      }
      // second layer, two steps of 16 input units: z1[t][8 q + j] is unit
      // u0 + (j & 3) + 8 (j >> 2) of this lane's half, which is the k index
      // 8 half + j of the B operand; the A operand takes the same units
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int u0 = 32 * mi + 16 * q + 4 * half;
        s16x8 bh[2], bl[2];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            float z = z1[t][8 * q + j] + b1[u0 + (j & 3) + 8 * (j >> 2)];
            z = z > 0.f ? z : 0.f;
            short hi, lo;
            split_bf16(z, hi, lo);
            bh[t][j] = hi;
            bl[t][j] = lo;
          }
        const int row = 32 * mo + col;
        const s16x8 ah = __builtin_shufflevector(*(const bf16x4*)&w2h[row][u0], *(const bf16x4*)&w2h[row][u0 + 8],
                                                 0, 1, 2, 3, 4, 5, 6, 7);
        const s16x8 al = __builtin_shufflevector(*(const bf16x4*)&w2l[row][u0], *(const bf16x4*)&w2l[row][u0 + 8],
                                                 0, 1, 2, 3, 4, 5, 6, 7);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          acc[t] = mfma16(ah, bh[t], acc[t]);
          acc[t] = mfma16(ah, bl[t], acc[t]);
          acc[t] = mfma16(al, bh[t], acc[t]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int u = 32 * mo + 8 * (r / 4) + 4 * half + (r % 4);
        const float z = acc[t][r] + b2[u];
        part[t] = fmaf(z > 0.f ? z : 0.f, w3[u], part[t]);
      }
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) part[t] += __shfl_xor(part[t], 32, 64);
  if (!live) return;
  const float y = b3 + part[half];
  if (a.score) a.score[i] = y;
  if (!a.key) return;
  const u32 kk = score_key(y);
  a.key[i] = kk;
  if (a.sample && (i & (a.sstep - 1)) == 0) a.sample[i / a.sstep] = kk;
}

__global__ void fill_u32(u32* p, u32 n, u32 v) {
  for (u32 i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// Threshold of the selection (one workgroup): the key of rank `at` (0 = the
// largest) among the n <= 1024 SEL_PER_LANE sampled keys, by a
// most-significant-digit-first radix select over four 8-bit digits.  Each lane
// loads its samples once, into registers.  The keys cluster (items alike in
// their features score alike), so for the top digit a wave first adds its
// lanes' digits to the LDS histogram one distinct digit at a time (one atomic
// per digit, for all lanes holding it), for at most LEADER_ROUNDS digits; the
// lanes left, and every lane for the lower digits (spread out, and only under
// the chosen prefix), add theirs one atomic per lane.  0 when every item is to
// be gathered.
constexpr int LEADER_ROUNDS = 2;
__global__ __launch_bounds__(1024) void select_threshold(const u32* smp, u32 n, u32 at, int all, u32* thr) {
  __shared__ u32 hist[256];
  __shared__ u32 prefix_s, at_s;
  const int lane = threadIdx.x & 63;
  if (all) {
    if (threadIdx.x == 0) *thr = 0;
    return;
  }
  u32 v[SEL_PER_LANE];
#pragma unroll
  for (u32 rd = 0; rd < SEL_PER_LANE; ++rd) {
    const u32 q = rd * blockDim.x + threadIdx.x;
    v[rd] = q < n ? smp[q] : 0u;
  }
  if (threadIdx.x == 0) { prefix_s = 0; at_s = at; }
  for (int shift = 24; shift >= 0; shift -= 8) {
    for (u32 q = threadIdx.x; q < 256; q += blockDim.x) hist[q] = 0;
    __syncthreads();
    const u32 prefix = prefix_s;
    const u32 hmask = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
#pragma unroll
    for (u32 rd = 0; rd < SEL_PER_LANE; ++rd) {
      const u32 q = rd * blockDim.x + threadIdx.x;
      bool pend = q < n && (v[rd] & hmask) == prefix;
      const u32 dg = (v[rd] >> shift) & 255u;
      for (int it = 0; it < (shift == 24 ? LEADER_ROUNDS : 0); ++it) {
        const u64 m = __ballot(pend);
        if (!m) break;
        const int leader = __ffsll((unsigned long long)m) - 1;
        const u32 ld = (u32)__shfl((int)dg, leader, 64);
        const u64 same = __ballot(pend && dg == ld);
        if (lane == leader) atomicAdd(&hist[ld], (u32)__popcll(same));
        if (pend && dg == ld) pend = false;
      }
      // digits still pending are spread out: one atomic per lane
      if (pend) atomicAdd(&hist[dg], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      // digits from the top: the one whose keys cover rank at_s.  Lane l holds
      // digits 4 l .. 4 l + 3; a suffix sum over the lanes gives the keys above
      // each lane's digits, and the lane whose range covers the rank (lane 0
      // when the rank is beyond every key) picks the digit.
      const u32 l = threadIdx.x;
      u32 h[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) h[j] = hist[4 * l + j];
      const u32 tot = h[0] + h[1] + h[2] + h[3];
      u32 incl = tot;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const u32 o = (u32)__shfl_down((int)incl, off, 64);
        if (l + off < 64) incl += o;
      }
      const u32 r = at_s, above = incl - tot;
      if ((r >= above && r < incl) || (l == 0 && r >= incl)) {
        u32 rr = r - above, d = 4 * l + 3;
        for (int j = 3; j > 0 && rr >= h[j]; --j, --d) rr -= h[j];
        prefix_s = prefix | (d << shift);
        at_s = rr;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *thr = prefix_s;
}

// Items with key >= *thr, gathered as entries key << 32 | ~item (a descending
// sort puts ties in ascending item order).  A workgroup collects the passing
// items of its SELECT_ITEMS items in LDS and claims their output slots with
// one atomic; slots beyond cap are dropped (cnt still counts them).
__global__ __launch_bounds__(256) void score_select(const u32* key, u32 n, const u32* thr_p, u32 cap, u32* cnt,
                                                    u64* out) {
  __shared__ u64 list[SELECT_ITEMS];
  __shared__ u32 nl, base;
  if (threadIdx.x == 0) nl = 0;
  __syncthreads();
  const u32 thr = *thr_p;
  const u32 i0 = blockIdx.x * SELECT_ITEMS;
  for (u32 q = threadIdx.x; q < SELECT_ITEMS; q += blockDim.x) {
    const u32 i = i0 + q;
    if (i >= n) break;
    const u32 kk = key[i];
    if (kk >= thr) list[atomicAdd(&nl, 1u)] = ((u64)kk << 32) | (u64)(~i);
  }
  __syncthreads();
  if (threadIdx.x == 0) base = nl ? atomicAdd(cnt, nl) : 0u;
  __syncthreads();
  for (u32 q = threadIdx.x; q < nl; q += blockDim.x)
    if (base + q < cap) out[base + q] = list[q];
}

// Sorts the gathered entries best first by rank: the entries are distinct
// (each holds its item), so an entry's position is the number of entries
// above it.  Every workgroup stages all entries in LDS; RANK_SPLIT lanes rank
// one entry together, each over a strided share of the entries, and add their
// counts by shuffles.  out[0] = entries gathered (may exceed the cap),
// out[1 ..] = the sorted entries.
constexpr u32 RANK_BLOCK = 256;
constexpr u32 RANK_SPLIT = 16;
constexpr u32 RANK_PER_BLOCK = RANK_BLOCK / RANK_SPLIT;
constexpr u32 RANK_UNROLL = 8;
static_assert(FRONT_CAP % (RANK_UNROLL * RANK_SPLIT) == 0, "the padded entries fit the LDS array");
__global__ __launch_bounds__(RANK_BLOCK) void sort_front(const u32* cnt, const u64* sel, u64* out) {
  __shared__ u64 s[FRONT_CAP];
  const u32 m = *cnt < FRONT_CAP ? *cnt : FRONT_CAP;
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = *cnt;
  if (blockIdx.x * RANK_PER_BLOCK >= m) return;
  // FRONT_CAP / RANK_BLOCK loads per lane, issued together; the entries are
  // padded with zeros (above no entry) to a multiple of RANK_UNROLL RANK_SPLIT
  const u32 mp = (m + RANK_UNROLL * RANK_SPLIT - 1) / (RANK_UNROLL * RANK_SPLIT) * (RANK_UNROLL * RANK_SPLIT);
#pragma unroll
  for (u32 r = 0; r < FRONT_CAP / RANK_BLOCK; ++r) {
    const u32 q = r * RANK_BLOCK + threadIdx.x;
    if (q < mp) s[q] = q < m ? sel[q] : 0ull;
  }
  __syncthreads();
  const u32 e_idx = blockIdx.x * RANK_PER_BLOCK + threadIdx.x / RANK_SPLIT;
  const u32 part = threadIdx.x % RANK_SPLIT;
  const u64 e = e_idx < m ? s[e_idx] : ~0ull;
  // RANK_UNROLL independent counts, so that that many LDS reads are in flight
  u32 above[RANK_UNROLL] = {};
  for (u32 q = part; q < mp; q += RANK_UNROLL * RANK_SPLIT) {
#pragma unroll
    for (u32 j = 0; j < RANK_UNROLL; ++j) above[j] += s[q + j * RANK_SPLIT] > e ? 1u : 0u;
  }
  u32 rank = 0;
#pragma unroll
  for (u32 j = 0; j < RANK_UNROLL; ++j) rank += above[j];
  // the RANK_SPLIT lanes of an entry are consecutive in the wave
#pragma unroll
  for (u32 off = RANK_SPLIT / 2; off > 0; off >>= 1) rank += (u32)__shfl_xor((int)rank, (int)off, 64);
  if (part == 0 && e_idx < m) out[1 + rank] = e;
}

}  // namespace

// The front's pinned staging, after the model parameters in the context's
// pinned buffer.
static u64* front_stage(gmk_ctx* c) { return (u64*)(c->h_stage + (MLP_NPARAM + 2 * IN + 1) / 2 * 2); }

namespace {
// First n items of the sorted front as one-vertex tuples (vertex IDs).
__global__ void front_tuples(const u64* sorted, u32 a_lo, u32 n, u32* tuples) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tuples[i] = a_lo + ~(u32)(sorted[1 + i] & 0xFFFFFFFFull);
}
}  // namespace

int gmk::learned_front(gmk_ctx* c, u32 a_lo, u32 a_hi, u32 k, u32 want, u32 pre_max, std::vector<u64>& front,
                       u32& pre, double& ms) {
  front.clear();
  pre = 0;
  ms = 0;
  const u32 n = a_hi - a_lo;
  if (!c->has_item_model || n == 0 || want == 0) return GMK_OK;
  want = std::min(want, std::min(n, FRONT_CAP / 2));
  hipStream_t s = c->stream;
  int r;
  // sample stride: a power of two, so that at most 1024 SEL_PER_LANE samples
  u32 sstep = SSTEP_MIN;
  while ((u64)(n + sstep - 1) / sstep > 1024ull * SEL_PER_LANE) sstep <<= 1;
  const u32 ns = (n + sstep - 1) / sstep;
  // device: parameters, then (u32) count, threshold, keys, samples; entries;
  // sorted front.  The parameters and the sorted front pass through the
  // context's pinned staging.
  if ((r = ensure(c, c->sched_p, (MLP_NPARAM + 2 * IN) * 4))) return r;
  if ((r = ensure(c, c->sched_hist, 16 + 4 * ((size_t)n + ns)))) return r;
  if ((r = ensure(c, c->sched_sel, (size_t)2 * (FRONT_CAP + 1) * 8))) return r;
  static_assert((size_t)(MLP_NPARAM + 2 * IN) * 4 + (FRONT_CAP + 1) * 8 <=
                    ((size_t)MLP_NPARAM + 2 * IN + GMK_STATIC_ITEMS) * 4,
                "the pinned staging holds the parameters and the sorted front");
  float* P = c->h_stage;
  u64* hfront = front_stage(c);
  mlp_pack(c->item_model, P);
  for (int q = 0; q < IN; ++q) { P[MLP_NPARAM + q] = c->item_model.mu[q]; P[MLP_NPARAM + IN + q] = c->item_model.sd[q]; }
  u32* cnt = (u32*)c->sched_hist.p;
  u32* thr = cnt + 1;
  u32* key = cnt + 4;
  u32* smp = key + n;
  u64* sel = (u64*)c->sched_sel.p;
  u64* sorted = sel + FRONT_CAP + 1;
  HIPCHK(c, hipMemcpyAsync(c->sched_p.p, P, (MLP_NPARAM + 2 * IN) * 4, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipEventRecord(c->ev2, s));
  hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, s, cnt, 4u, 0u);
  ItemArgs a{c->d_off, c->d_adj, c->d_deg, a_lo, n, (a_lo == 0 ? c->n - c->n_left : c->n_left), (float)k,
             (const float*)c->sched_p.p, nullptr, key, smp, nullptr, sstep};
  hipLaunchKernelGGL(item_score, dim3((n + 255) / 256), dim3(256), 0, s, a);
  // Threshold: the sampled key ranked at about 1.5 `want` items, so that the
  // gathered set holds the best `want` items with room for the sample's error
  // and stays under FRONT_CAP; every item when there are at most 1.5 `want`.
  const u64 span = (u64)want * 3 / 2;
  const u32 at = (u32)std::min<u64>(ns - 1, (span + sstep - 1) / sstep);
  const bool all = n <= span;
  hipLaunchKernelGGL(select_threshold, dim3(1), dim3(1024), 0, s, (const u32*)smp, ns, at, all ? 1 : 0, thr);
  hipLaunchKernelGGL(score_select, dim3((n + SELECT_ITEMS - 1) / SELECT_ITEMS), dim3(256), 0, s, (const u32*)key, n,
                     (const u32*)thr, FRONT_CAP, cnt, sel);
  hipLaunchKernelGGL(sort_front, dim3(FRONT_CAP / RANK_PER_BLOCK), dim3(RANK_BLOCK), 0, s, (const u32*)cnt,
                     (const u64*)sel, sorted);
  // The front holds at least every item (all) or the at + 1 sampled items at
  // or above the threshold, up to FRONT_CAP: when that is pre_max or more, its
  // first pre_max items become the first launch's tuples on the device.
  const u64 front_min = std::min<u64>(all ? n : (u64)at + 1, FRONT_CAP);
  if (pre_max && front_min >= pre_max) {
    hipLaunchKernelGGL(front_tuples, dim3((pre_max + 255) / 256), dim3(256), 0, s, (const u64*)sorted, a_lo, pre_max,
                       (u32*)c->tuples.p);
    pre = pre_max;
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev3, s));
  HIPCHK(c, hipMemcpyAsync(hfront, sorted, (size_t)(FRONT_CAP + 1) * 8, hipMemcpyDeviceToHost, s));
  if (pre) return GMK_OK;
  HIPCHK(c, hipStreamSynchronize(s));
  learned_front_finish(c, front, ms);
  return GMK_OK;
}

void gmk::learned_front_finish(gmk_ctx* c, std::vector<u64>& front, double& ms) {
  const u64* hfront = front_stage(c);
  // entries gathered beyond FRONT_CAP (a rare sampling error) are left out of
  // the front; their items follow in ID order like every item outside it
  const u64 m = std::min<u64>(hfront[0], FRONT_CAP);
  front.assign(hfront + 1, hfront + 1 + m);
  float t = 0;
  ms = hipEventElapsedTime(&t, c->ev2, c->ev3) == hipSuccess ? t : 0.0;
}

u32 gmk::LearnedOrder::next(u32 a_lo) { return a_lo + ~(u32)(cand[pos++] & 0xFFFFFFFFu); }

// Features (predict = false) or predictions of the item model (predict = true)
// of items first .. first + count - 1, by the scheduling kernel itself.
static int item_eval(gmk_ctx* c, uint32_t first, uint32_t count, uint32_t k, float* out, bool predict) {
  if (!c || (!out && count)) return GMK_EINVAL;
  if (c->n == 0) { c->err = "no graph loaded"; return GMK_ENOGRAPH; }
  if (predict && !c->has_item_model) { c->err = "no item model set"; return GMK_EINVAL; }
  if (count == 0) return GMK_OK;
  const bool left = first < c->n_left;
  const u64 end = (u64)first + count;
  if (end > (left ? c->n_left : c->n)) { c->err = "items must lie on one side of the graph"; return GMK_EINVAL; }
  HIPCHK(c, hipSetDevice(c->device));
  const size_t width = predict ? 1 : IN;
  DeviceVec f, p;
  int r;
  if ((r = ensure(c, f, (size_t)count * width * 4))) return r;
  std::vector<float> P(MLP_NPARAM + 2 * IN);
  if (predict) {
    if ((r = ensure(c, p, P.size() * 4))) { (void)hipFree(f.p); return r; }
    mlp_pack(c->item_model, P.data());
    for (int q = 0; q < IN; ++q) { P[MLP_NPARAM + q] = c->item_model.mu[q]; P[MLP_NPARAM + IN + q] = c->item_model.sd[q]; }
  }
  ItemArgs a{c->d_off, c->d_adj, c->d_deg, first, count, left ? c->n - c->n_left : c->n_left, (float)k,
             predict ? (const float*)p.p : nullptr, predict ? nullptr : (float*)f.p, nullptr, nullptr,
             predict ? (float*)f.p : nullptr};
  hipError_t e = hipSuccess;
  if (predict) e = hipMemcpyAsync(p.p, P.data(), P.size() * 4, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(item_score, dim3((count + 255) / 256), dim3(256), 0, c->stream, a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, f.p, (size_t)count * width * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(f.p);
  if (p.p) (void)hipFree(p.p);
  if (e != hipSuccess) { c->err = std::string("item kernel: ") + hipGetErrorString(e); return GMK_EDEVICE; }
  return GMK_OK;
}

// ---------------------------------------------------------------------------
// Predictions with no allocation (gmk_item_predict_static).  Its scratch is a
// file-scope device array: the model parameters (gmk::mlp_pack order, then mu
// and sd), then the scores of up to GMK_STATIC_ITEMS items.  In host code the
// array's name is not a device address: gmk_create takes this device's address
// of it once (static_scratch, with the device current) and keeps it in the
// context; host code uses only that address.
// ---------------------------------------------------------------------------
constexpr size_t SCRATCH_PARAMS = (size_t)MLP_NPARAM + 2 * IN;
constexpr size_t SCRATCH_FLOATS = SCRATCH_PARAMS + GMK_STATIC_ITEMS;
static __device__ __align__(256) float scratch[SCRATCH_FLOATS];

int gmk::static_scratch(float** ptr, size_t* floats) {
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GMK_EDEVICE;
  *ptr = (float*)p;
  *floats = SCRATCH_FLOATS;
  return GMK_OK;
}

extern "C" int gmk_item_predict_static(gmk_ctx* c, uint32_t first, uint32_t count, uint32_t k, float* out) {
  if (!c || (!out && count)) return GMK_EINVAL;
  if (c->n == 0) { c->err = "no graph loaded"; return GMK_ENOGRAPH; }
  if (!c->has_item_model) { c->err = "no item model set"; return GMK_EINVAL; }
  if (count > GMK_STATIC_ITEMS) {
    c->err = "gmk_item_predict_static: " + std::to_string(count) + " items asked, at most " +
             std::to_string(GMK_STATIC_ITEMS) + " per call (the static scratch array's size)";
    return GMK_EINVAL;
  }
  if (count == 0) return GMK_OK;
  const bool left = first < c->n_left;
  if ((u64)first + count > (left ? c->n_left : c->n)) { c->err = "items must lie on one side of the graph"; return GMK_EINVAL; }
  if (!c->scratch || !c->h_stage) { c->err = "static scratch not set up"; return GMK_EDEVICE; }
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  // the parameters go through the context's pinned staging (no allocation)
  mlp_pack(c->item_model, c->h_stage);
  for (int q = 0; q < IN; ++q) { c->h_stage[MLP_NPARAM + q] = c->item_model.mu[q]; c->h_stage[MLP_NPARAM + IN + q] = c->item_model.sd[q]; }
  HIPCHK(c, hipMemcpyAsync(c->scratch, c->h_stage, SCRATCH_PARAMS * 4, hipMemcpyHostToDevice, s));
  float* score = c->scratch + SCRATCH_PARAMS;
  ItemArgs a{c->d_off, c->d_adj, c->d_deg, first, count, left ? c->n - c->n_left : c->n_left, (float)k, c->scratch, nullptr,
             nullptr, nullptr, score};
  hipLaunchKernelGGL(item_score, dim3((count + 255) / 256), dim3(256), 0, s, a);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->h_stage + SCRATCH_PARAMS, score, (size_t)count * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  memcpy(out, c->h_stage + SCRATCH_PARAMS, (size_t)count * 4);
  return GMK_OK;
}

extern "C" int gmk_item_features(gmk_ctx* c, uint32_t first, uint32_t count, uint32_t k, float* out) {
  return item_eval(c, first, count, k, out, false);
}

extern "C" int gmk_item_predict(gmk_ctx* c, uint32_t first, uint32_t count, uint32_t k, float* out) {
  return item_eval(c, first, count, k, out, true);
}

extern "C" int gmk_set_item_model(gmk_ctx* c, const gmk_mlp* model) {
  if (!c) return GMK_EINVAL;
  c->has_item_model = model != nullptr;
  if (model) c->item_model = *model;
  return GMK_OK;
}
