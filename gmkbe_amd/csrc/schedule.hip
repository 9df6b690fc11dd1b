// Learned schedule of single-graph runs.  A work item is a one-vertex tuple of
// the thin enumerator (a vertex u of side A with its families: u with N(u) and
// up to k vertices it misses).  The item cost model (the FP32 MLP of mlp.hip,
// trained on the yield and device time the method logs per item, see
// gmk_probe_items) predicts log(results per ms of device lane time) from the
// item's features; a limited run takes the items with the best prediction
// first.  The model only reorders work: every item is still enumerated
// exactly, and a run without a model keeps the degree order of the IDs.
//
// Device work: one lane per item computes its features and the forward pass
// (item_score), every 64th score is sampled to set the threshold of the best
// `want` items, and the items above it are gathered (score_select).  The host
// sorts only the gathered items.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gmk_host.h"

using namespace gmk;

namespace {

constexpr int IN = GMK_MLP_IN, H = GMK_MLP_HID;
constexpr u32 SAMPLE = 32;           // neighbours sampled per item for the degree moments
constexpr u32 SSTEP = 64;           // every SSTEP-th item's key is sampled for the selection threshold
constexpr u32 SEL_MAX = 1u << 17;    // gathered items, at most

struct ItemArgs {
  const u32* off;
  const u32* adj;
  u32 first, n;     // items: vertices first .. first + n - 1
  u32 nb;           // vertices on side B
  float k;
  const float* P;   // model parameters (gmk::mlp_pack order), then mu[IN], sd[IN]; null: features only
  float* feat;      // n x IN features, or null
  u32* key;         // order-preserving score keys, or null
  u32* sample;      // key of every SSTEP-th item, or null
  float* score;     // predictions, or null
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
  // mean and variance by Welford's update (no cancellation when the degrees agree)
  float mean = 0.f, m2 = 0.f, mx = 0.f;
  u32 mn = 0xFFFFFFFFu;
  double hop = 0;
  for (u32 i = 0; i < ns; ++i) {
    const u32 e = ns == d ? i : (u32)(((u64)i * d) / ns);
    const u32 v = a.adj[o + e];
    const u32 dv = a.off[v + 1] - a.off[v];
    const float l = log1pf((float)dv);
    const float delta = l - mean;
    mean += delta / (float)(i + 1);
    m2 += delta * (l - mean);
    mx = l > mx ? l : mx;
    mn = dv < mn ? dv : mn;
    hop += (double)(dv - 1);
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

// One lane per item: features, then (with a model) the forward pass with the
// hidden layer in registers and the weights in LDS (w2 transposed so a lane
// reads one output's weights contiguously).
__global__ __launch_bounds__(256) void item_score(ItemArgs a) {
  __shared__ float w1[IN * H], b1[H], w2t[H * H], b2[H], w3[H], mu[IN], sd[IN];
  __shared__ float b3;
  if (a.P) {
    const float* P = a.P;
    constexpr int OW1 = 0, OB1 = OW1 + IN * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
    for (int i = threadIdx.x; i < IN * H; i += blockDim.x) w1[i] = P[OW1 + i];
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) w2t[(i % H) * H + i / H] = P[OW2 + i];
    for (int i = threadIdx.x; i < H; i += blockDim.x) { b1[i] = P[OB1 + i]; b2[i] = P[OB2 + i]; w3[i] = P[OW3 + i]; }
    if (threadIdx.x < IN) { mu[threadIdx.x] = P[OB3 + 1 + threadIdx.x]; sd[threadIdx.x] = P[OB3 + 1 + IN + threadIdx.x]; }
    if (threadIdx.x == 0) b3 = P[OB3];
    __syncthreads();
  }
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float x[IN];
  item_features(a, a.first + i, x);
  if (a.feat) {
#pragma unroll
    for (int q = 0; q < IN; ++q) a.feat[(u64)i * IN + q] = x[q];
  }
  if (!a.P) return;
#pragma unroll
  for (int q = 0; q < IN; ++q) x[q] = (x[q] - mu[q]) / sd[q];
  float h1[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float z = b1[j];
#pragma unroll
    for (int q = 0; q < IN; ++q) z = fmaf(x[q], w1[q * H + j], z);
    h1[j] = z > 0.f ? z : 0.f;
  }
  float y = b3;
  for (int j = 0; j < H; ++j) {
    float z = b2[j];
#pragma unroll
    for (int q = 0; q < H; ++q) z = fmaf(h1[q], w2t[j * H + q], z);
    y = fmaf(z > 0.f ? z : 0.f, w3[j], y);
  }
  if (a.score) a.score[i] = y;
  if (!a.key) return;
  const u32 kk = score_key(y);
  a.key[i] = kk;
  if (a.sample && i % SSTEP == 0) a.sample[i / SSTEP] = kk;
}

// Items with key >= thr, gathered with one counter atomic per wave (a wave's
// passing lanes take consecutive slots).  Entries: key << 32 | ~item, so a
// descending sort puts ties in ascending item order.
__global__ __launch_bounds__(256) void score_select(const u32* key, u32 n, u32 thr, u32 cap, u32* cnt, u64* out) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const u32 kk = i < n ? key[i] : 0u;
  const bool pass = i < n && kk >= thr;
  const u64 m = __ballot(pass);
  if (!m) return;
  const int leader = __ffsll((unsigned long long)m) - 1;
  u32 base = 0;
  if (lane == leader) base = atomicAdd(cnt, (u32)__popcll(m));
  base = (u32)__shfl((int)base, leader, 64);
  const u32 at = base + (u32)__popcll(m & ((1ull << lane) - 1ull));
  if (pass && at < cap) out[at] = ((u64)kk << 32) | (u64)(~i);
}

}  // namespace

int gmk::learned_front(gmk_ctx* c, u32 a_lo, u32 a_hi, u32 k, u32 want, std::vector<u64>& front) {
  front.clear();
  const u32 n = a_hi - a_lo;
  if (!c->has_item_model || n == 0 || want == 0) return GMK_OK;
  want = std::min(want, std::min(n, SEL_MAX / 4));
  hipStream_t s = c->stream;
  int r;
  const u32 ns = (n + SSTEP - 1) / SSTEP;
  if ((r = ensure(c, c->sched_p, (MLP_NPARAM + 2 * IN) * 4))) return r;
  if ((r = ensure(c, c->sched_hist, 16 + 4 * ((size_t)n + ns)))) return r;
  if ((r = ensure(c, c->sched_sel, (size_t)SEL_MAX * 8))) return r;
  std::vector<float> P(MLP_NPARAM + 2 * IN);
  mlp_pack(c->item_model, P.data());
  for (int q = 0; q < IN; ++q) { P[MLP_NPARAM + q] = c->item_model.mu[q]; P[MLP_NPARAM + IN + q] = c->item_model.sd[q]; }
  u32* cnt = (u32*)c->sched_hist.p;
  u32* key = cnt + 4;
  u32* smp = key + n;
  HIPCHK(c, hipMemcpyAsync(c->sched_p.p, P.data(), P.size() * 4, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemsetAsync(cnt, 0, 16, s));
  ItemArgs a{c->d_off, c->d_adj, a_lo, n, (a_lo == 0 ? c->n - c->n_left : c->n_left), (float)k,
             (const float*)c->sched_p.p, nullptr, key, smp, nullptr};
  const dim3 g((n + 255) / 256), b(256);
  hipLaunchKernelGGL(item_score, g, b, 0, s, a);
  HIPCHK(c, hipGetLastError());
  // Threshold from the sampled keys: the key of the sample ranked at about
  // twice `want` items, so the gathered set holds the best `want` items with
  // room for the sample's error (it is gathered whole up to SEL_MAX).
  std::vector<u32> h(ns);
  HIPCHK(c, hipMemcpyAsync(h.data(), smp, (size_t)ns * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  const size_t at = std::min<size_t>(ns - 1, (2ull * want + SSTEP - 1) / SSTEP);
  std::nth_element(h.begin(), h.begin() + at, h.end(), std::greater<u32>());
  const u32 thr = n <= 2ull * want ? 0u : h[at];
  hipLaunchKernelGGL(score_select, g, b, 0, s, (const u32*)key, n, thr, SEL_MAX, cnt, (u64*)c->sched_sel.p);
  HIPCHK(c, hipGetLastError());
  u32 m = 0;
  HIPCHK(c, hipMemcpyAsync(&m, cnt, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  m = std::min(m, SEL_MAX);
  std::vector<u64> sel(m);
  if (m) {
    HIPCHK(c, hipMemcpyAsync(sel.data(), c->sched_sel.p, (size_t)m * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
  }
  front.swap(sel);
  return GMK_OK;
}

u32 gmk::LearnedOrder::next(u32 a_lo) {
  if (pos == sorted) {
    // bring the next best STEP entries to the front of the rest, sorted
    const size_t end = std::min(cand.size(), pos + STEP);
    std::nth_element(cand.begin() + pos, cand.begin() + (end - 1), cand.end(), std::greater<u64>());
    std::sort(cand.begin() + pos, cand.begin() + end, std::greater<u64>());
    sorted = end;
  }
  return a_lo + ~(u32)(cand[pos++] & 0xFFFFFFFFu);
}

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
  ItemArgs a{c->d_off, c->d_adj, first, count, left ? c->n - c->n_left : c->n_left, (float)k,
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
  ItemArgs a{c->d_off, c->d_adj, first, count, left ? c->n - c->n_left : c->n_left, (float)k, c->scratch, nullptr,
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
