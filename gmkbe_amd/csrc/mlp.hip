// FP32 MLP cost model: forward, backward, weight gradients and Adam, as plain
// FMA loops (one thread per sample or per parameter).  The network is tiny
// (8 -> 64 -> 64 -> 1); this is scheduling support, not a matrix workload.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "gmk_host.h"

using namespace gmk;

namespace {

constexpr int IN = GMK_MLP_IN, H = GMK_MLP_HID;
constexpr int NPARAM = gmk::MLP_NPARAM;  // w1 b1 w2 b2 w3 b3
// parameter offsets inside the flat vector (same order as gmk_mlp)
constexpr int OW1 = 0, OB1 = OW1 + IN * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;

// Forward pass, one thread per sample: standardised input x (n x IN).
// Keeps z1, z2 (pre-activations) and y for the backward pass.
__global__ __launch_bounds__(256) void mlp_forward(const float* __restrict__ P, const float* __restrict__ X, uint32_t n,
                                                   float* __restrict__ Z1, float* __restrict__ Z2, float* __restrict__ Y) {
  __shared__ float w2[H * H];
  for (int i = threadIdx.x; i < H * H; i += blockDim.x) w2[i] = P[OW2 + i];
  __syncthreads();
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[IN];
  for (int i = 0; i < IN; ++i) x[i] = X[(size_t)s * IN + i];
  float a1[H];
  for (int j = 0; j < H; ++j) {
    float z = P[OB1 + j];
    for (int i = 0; i < IN; ++i) z = fmaf(x[i], P[OW1 + i * H + j], z);
    if (Z1) Z1[(size_t)s * H + j] = z;
    a1[j] = z > 0.f ? z : 0.f;
  }
  float y = P[OB3];
  for (int j = 0; j < H; ++j) {
    float z = P[OB2 + j];
    for (int i = 0; i < H; ++i) z = fmaf(a1[i], w2[i * H + j], z);
    if (Z2) Z2[(size_t)s * H + j] = z;
    y = fmaf(z > 0.f ? z : 0.f, P[OW3 + j], y);
  }
  Y[s] = y;
}

// Backward pass per sample for loss = mean (y - t)^2: writes dZ2, dZ1 and dy.
__global__ __launch_bounds__(256) void mlp_backward(const float* __restrict__ P, const float* __restrict__ T, uint32_t n,
                                                    const float* __restrict__ Z1, const float* __restrict__ Z2,
                                                    const float* __restrict__ Y, float* __restrict__ DZ1,
                                                    float* __restrict__ DZ2, float* __restrict__ DY) {
  __shared__ float w2[H * H];
  for (int i = threadIdx.x; i < H * H; i += blockDim.x) w2[i] = P[OW2 + i];
  __syncthreads();
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float dy = 2.f * (Y[s] - T[s]) / (float)n;
  DY[s] = dy;
  float dz2[H];
  for (int j = 0; j < H; ++j) {
    const float z = Z2[(size_t)s * H + j];
    dz2[j] = z > 0.f ? dy * P[OW3 + j] : 0.f;
    DZ2[(size_t)s * H + j] = dz2[j];
  }
  for (int i = 0; i < H; ++i) {
    float da = 0.f;
    for (int j = 0; j < H; ++j) da = fmaf(w2[i * H + j], dz2[j], da);
    DZ1[(size_t)s * H + i] = Z1[(size_t)s * H + i] > 0.f ? da : 0.f;
  }
}

// Gradient of every parameter: one thread per (parameter, slice of the
// samples) writes that slice's sum; mlp_wgrad_reduce adds the WG_SLICES
// partial sums in a fixed order (deterministic).
constexpr int WG_SLICES = 64;
__global__ __launch_bounds__(256) void mlp_wgrad(const float* __restrict__ X, uint32_t n_all,
                                                 const float* __restrict__ Z1, const float* __restrict__ Z2,
                                                 const float* __restrict__ DZ1, const float* __restrict__ DZ2,
                                                 const float* __restrict__ DY, float* __restrict__ Gp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const uint32_t s0 = (uint32_t)(((uint64_t)n_all * blockIdx.y) / WG_SLICES);
  const uint32_t n = (uint32_t)(((uint64_t)n_all * (blockIdx.y + 1)) / WG_SLICES);
  float* G = Gp + (size_t)blockIdx.y * NPARAM;
  float acc = 0.f;
  if (p < OB1) {  // w1[i][j] = sum x_i dz1_j
    const int i = p / H, j = p % H;
    for (uint32_t s = s0; s < n; ++s) acc = fmaf(X[(size_t)s * IN + i], DZ1[(size_t)s * H + j], acc);
  } else if (p < OW2) {  // b1
    const int j = p - OB1;
    for (uint32_t s = s0; s < n; ++s) acc += DZ1[(size_t)s * H + j];
  } else if (p < OB2) {  // w2[i][j] = sum relu(z1_i) dz2_j
    const int q = p - OW2, i = q / H, j = q % H;
    for (uint32_t s = s0; s < n; ++s) {
      const float a = Z1[(size_t)s * H + i];
      if (a > 0.f) acc = fmaf(a, DZ2[(size_t)s * H + j], acc);
    }
  } else if (p < OW3) {  // b2
    const int j = p - OB2;
    for (uint32_t s = s0; s < n; ++s) acc += DZ2[(size_t)s * H + j];
  } else if (p < OB3) {  // w3[j] = sum relu(z2_j) dy
    const int j = p - OW3;
    for (uint32_t s = s0; s < n; ++s) {
      const float a = Z2[(size_t)s * H + j];
      if (a > 0.f) acc = fmaf(a, DY[s], acc);
    }
  } else {  // b3
    for (uint32_t s = s0; s < n; ++s) acc += DY[s];
  }
  G[p] = acc;
}

__global__ __launch_bounds__(256) void mlp_wgrad_reduce(const float* __restrict__ Gp, float* __restrict__ G) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float acc = 0.f;
  for (int q = 0; q < WG_SLICES; ++q) acc += Gp[(size_t)q * NPARAM + p];
  G[p] = acc;
}

__global__ __launch_bounds__(256) void mlp_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                                                float* __restrict__ V, float lr, float b1t, float b2t) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float g = G[p];
  const float m = 0.9f * M[p] + 0.1f * g;
  const float v = 0.999f * V[p] + 0.001f * g * g;
  M[p] = m;
  V[p] = v;
  const float mh = m / (1.f - b1t), vh = v / (1.f - b2t);
  P[p] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

}  // namespace

void gmk::mlp_pack(const gmk_mlp& m, float* P) {
  memcpy(P + OW1, m.w1, sizeof m.w1);
  memcpy(P + OB1, m.b1, sizeof m.b1);
  memcpy(P + OW2, m.w2, sizeof m.w2);
  memcpy(P + OB2, m.b2, sizeof m.b2);
  memcpy(P + OW3, m.w3, sizeof m.w3);
  P[OB3] = m.b3;
}

namespace {
void pack(const gmk_mlp& m, float* P) { gmk::mlp_pack(m, P); }
void unpack(const float* P, gmk_mlp& m) {
  memcpy(m.w1, P + OW1, sizeof m.w1);
  memcpy(m.b1, P + OB1, sizeof m.b1);
  memcpy(m.w2, P + OW2, sizeof m.w2);
  memcpy(m.b2, P + OB2, sizeof m.b2);
  memcpy(m.w3, P + OW3, sizeof m.w3);
  m.b3 = P[OB3];
}

struct MlpBuffers {
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr, *X = nullptr, *T = nullptr, *Y = nullptr;
  float *Z1 = nullptr, *Z2 = nullptr, *DZ1 = nullptr, *DZ2 = nullptr, *DY = nullptr, *GP = nullptr;
  ~MlpBuffers() {
    for (float* p : {P, G, M, V, X, T, Y, Z1, Z2, DZ1, DZ2, DY, GP})
      if (p) (void)hipFree(p);
  }
};

int alloc_mlp(gmk_ctx* c, MlpBuffers& b, uint32_t n) {
  auto al = [&](float** p, size_t cnt) -> hipError_t { return hipMalloc((void**)p, std::max<size_t>(1, cnt) * 4); };
  HIPCHK(c, al(&b.P, NPARAM));
  HIPCHK(c, al(&b.G, NPARAM));
  HIPCHK(c, al(&b.M, NPARAM));
  HIPCHK(c, al(&b.V, NPARAM));
  HIPCHK(c, al(&b.X, (size_t)n * IN));
  HIPCHK(c, al(&b.T, n));
  HIPCHK(c, al(&b.Y, n));
  HIPCHK(c, al(&b.Z1, (size_t)n * H));
  HIPCHK(c, al(&b.Z2, (size_t)n * H));
  HIPCHK(c, al(&b.DZ1, (size_t)n * H));
  HIPCHK(c, al(&b.DZ2, (size_t)n * H));
  HIPCHK(c, al(&b.DY, n));
  HIPCHK(c, al(&b.GP, (size_t)WG_SLICES * NPARAM));
  return GMK_OK;
}

void standardise(const gmk_mlp& m, const float* X, uint32_t n, std::vector<float>& out) {
  out.resize((size_t)n * IN);
  for (uint32_t s = 0; s < n; ++s)
    for (int i = 0; i < IN; ++i) out[(size_t)s * IN + i] = (X[(size_t)s * IN + i] - m.mu[i]) / m.sd[i];
}

// forward (+ optional backward and gradients) on the device
int run_fb(gmk_ctx* c, MlpBuffers& b, uint32_t n, bool grad) {
  const dim3 blk(256), gs((n + 255) / 256), gp((NPARAM + 255) / 256);
  hipLaunchKernelGGL(mlp_forward, gs, blk, 0, c->stream, b.P, b.X, n, b.Z1, b.Z2, b.Y);
  HIPCHK(c, hipGetLastError());
  if (grad) {
    hipLaunchKernelGGL(mlp_backward, gs, blk, 0, c->stream, b.P, b.T, n, b.Z1, b.Z2, b.Y, b.DZ1, b.DZ2, b.DY);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(mlp_wgrad, dim3(gp.x, WG_SLICES), blk, 0, c->stream, b.X, n, b.Z1, b.Z2, b.DZ1, b.DZ2, b.DY,
                       b.GP);
    hipLaunchKernelGGL(mlp_wgrad_reduce, gp, blk, 0, c->stream, b.GP, b.G);
    HIPCHK(c, hipGetLastError());
  }
  return GMK_OK;
}

}  // namespace

extern "C" int gmk_graph_features(const gmk_graph_desc* g, float out[GMK_MLP_IN]) {
  if (!g || !out || !g->off) return GMK_EINVAL;
  const double nl = g->n_left, nr = (double)(g->n - g->n_left);
  const double e = g->off[g->n] / 2.0;
  uint32_t maxd = 0;
  double two_hop = 0;  // paths of length two: sum over vertices of C(deg, 2)
  for (uint32_t v = 0; v < g->n; ++v) {
    const uint32_t d = g->off[v + 1] - g->off[v];
    maxd = std::max(maxd, d);
    two_hop += 0.5 * (double)d * (double)(d > 0 ? d - 1 : 0);
  }
  out[0] = (float)std::log1p(nl);
  out[1] = (float)std::log1p(nr);
  out[2] = (float)std::log1p(e);
  out[3] = (float)std::log1p(nl > 0 ? e / nl : 0);
  out[4] = (float)std::log1p(nr > 0 ? e / nr : 0);
  out[5] = (float)std::log1p((double)maxd);
  out[6] = (float)std::log1p(two_hop);
  out[7] = (float)g->k;
  return GMK_OK;
}

extern "C" int gmk_mlp_train(gmk_ctx* c, const float* X, const float* y, uint32_t n, uint32_t epochs, float lr,
                             uint32_t seed, gmk_mlp* model, float* loss) {
  if (!c || !X || !y || !model || n == 0) return GMK_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  gmk_mlp m{};
  for (int i = 0; i < IN; ++i) {
    double mu = 0, sq = 0;
    for (uint32_t s = 0; s < n; ++s) mu += X[(size_t)s * IN + i];
    mu /= n;
    for (uint32_t s = 0; s < n; ++s) sq += (X[(size_t)s * IN + i] - mu) * (X[(size_t)s * IN + i] - mu);
    const double sd = std::sqrt(sq / n);
    m.mu[i] = (float)mu;
    m.sd[i] = sd > 1e-6 ? (float)sd : 1.f;
  }
  // He initialisation from a seeded generator
  std::mt19937 rng(seed);
  std::normal_distribution<float> n1(0.f, std::sqrt(2.f / IN)), n2(0.f, std::sqrt(2.f / H)), n3(0.f, std::sqrt(1.f / H));
  for (float& w : m.w1) w = n1(rng);
  for (float& w : m.w2) w = n2(rng);
  for (float& w : m.w3) w = n3(rng);
  double ymean = 0;
  for (uint32_t s = 0; s < n; ++s) ymean += y[s];
  m.b3 = (float)(ymean / n);
  // Early stopping: with 20 or more samples a seeded fifth is held out; every 5
  // epochs the held-out error is measured and the best parameters are kept.
  std::vector<uint32_t> perm(n);
  for (uint32_t s = 0; s < n; ++s) perm[s] = s;
  const uint32_t nv = n >= 20 ? n / 5 : 0, nt = n - nv;
  if (nv) std::shuffle(perm.begin(), perm.end(), std::mt19937(seed ^ 0x5EEDu));
  std::vector<float> xs_all;
  standardise(m, X, n, xs_all);
  std::vector<float> xt((size_t)nt * IN), yt(nt), xv((size_t)nv * IN), yv(nv);
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t src = perm[s];
    float* dx = s < nt ? &xt[(size_t)s * IN] : &xv[(size_t)(s - nt) * IN];
    std::copy(&xs_all[(size_t)src * IN], &xs_all[(size_t)src * IN] + IN, dx);
    (s < nt ? yt[s] : yv[s - nt]) = y[src];
  }
  MlpBuffers b, bv;
  int r;
  if ((r = alloc_mlp(c, b, nt))) return r;
  if (nv && (r = alloc_mlp(c, bv, nv))) return r;
  std::vector<float> P(NPARAM), best(NPARAM);
  pack(m, P.data());
  best = P;
  HIPCHK(c, hipMemcpyAsync(b.P, P.data(), NPARAM * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(b.M, 0, NPARAM * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(b.V, 0, NPARAM * 4, c->stream));
  HIPCHK(c, hipMemcpyAsync(b.X, xt.data(), (size_t)nt * IN * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(b.T, yt.data(), (size_t)nt * 4, hipMemcpyHostToDevice, c->stream));
  if (nv) HIPCHK(c, hipMemcpyAsync(bv.X, xv.data(), (size_t)nv * IN * 4, hipMemcpyHostToDevice, c->stream));
  double best_err = 1e300;
  std::vector<float> pv(nv);
  auto score = [&]() -> int {  // held-out mean squared error of the current parameters
    hipLaunchKernelGGL(mlp_forward, dim3((nv + 255) / 256), dim3(256), 0, c->stream, b.P, bv.X, nv, bv.Z1, bv.Z2,
                       bv.Y);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(pv.data(), bv.Y, (size_t)nv * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double err = 0;
    for (uint32_t s = 0; s < nv; ++s) err += (double)(pv[s] - yv[s]) * (pv[s] - yv[s]);
    if (err < best_err) {
      best_err = err;
      HIPCHK(c, hipMemcpyAsync(best.data(), b.P, NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return GMK_OK;
  };
  if (nv && (r = score())) return r;
  float b1t = 1.f, b2t = 1.f;
  for (uint32_t e = 0; e < epochs; ++e) {
    if ((r = run_fb(c, b, nt, true))) return r;
    b1t *= 0.9f;
    b2t *= 0.999f;
    hipLaunchKernelGGL(mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, c->stream, b.P, b.G, b.M, b.V, lr, b1t, b2t);
    HIPCHK(c, hipGetLastError());
    if (nv && (e % 5 == 4 || e + 1 == epochs) && (r = score())) return r;
  }
  if (nv) HIPCHK(c, hipMemcpyAsync(b.P, best.data(), NPARAM * 4, hipMemcpyHostToDevice, c->stream));
  // final loss over every sample
  MlpBuffers ba;
  if ((r = alloc_mlp(c, ba, n))) return r;
  HIPCHK(c, hipMemcpyAsync(ba.X, xs_all.data(), (size_t)n * IN * 4, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(mlp_forward, dim3((n + 255) / 256), dim3(256), 0, c->stream, b.P, ba.X, n, ba.Z1, ba.Z2, ba.Y);
  HIPCHK(c, hipGetLastError());
  std::vector<float> yh(n);
  HIPCHK(c, hipMemcpyAsync(P.data(), b.P, NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(yh.data(), ba.Y, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  unpack(P.data(), m);
  if (loss) {
    double l = 0;
    for (uint32_t s = 0; s < n; ++s) l += (double)(yh[s] - y[s]) * (yh[s] - y[s]);
    *loss = (float)(l / n);
  }
  *model = m;
  return GMK_OK;
}

extern "C" int gmk_mlp_predict(gmk_ctx* c, const gmk_mlp* model, const float* X, uint32_t n, float* y) {
  if (!c || !model || !X || !y) return GMK_EINVAL;
  if (n == 0) return GMK_OK;
  HIPCHK(c, hipSetDevice(c->device));
  MlpBuffers b;
  int r;
  if ((r = alloc_mlp(c, b, n))) return r;
  std::vector<float> P(NPARAM), xs;
  pack(*model, P.data());
  standardise(*model, X, n, xs);
  HIPCHK(c, hipMemcpyAsync(b.P, P.data(), NPARAM * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(b.X, xs.data(), (size_t)n * IN * 4, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(mlp_forward, dim3((n + 255) / 256), dim3(256), 0, c->stream, b.P, b.X, n, (float*)nullptr,
                     (float*)nullptr, b.Y);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(y, b.Y, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GMK_OK;
}

extern "C" int gmk_mlp_grad(gmk_ctx* c, const gmk_mlp* model, const float* X, const float* y, uint32_t n, gmk_mlp* grad,
                            float* loss) {
  if (!c || !model || !X || !y || !grad || n == 0) return GMK_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  MlpBuffers b;
  int r;
  if ((r = alloc_mlp(c, b, n))) return r;
  std::vector<float> P(NPARAM), xs;
  pack(*model, P.data());
  standardise(*model, X, n, xs);
  HIPCHK(c, hipMemcpyAsync(b.P, P.data(), NPARAM * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(b.X, xs.data(), (size_t)n * IN * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(b.T, y, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  if ((r = run_fb(c, b, n, true))) return r;
  std::vector<float> G(NPARAM), yh(n);
  HIPCHK(c, hipMemcpyAsync(G.data(), b.G, NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(yh.data(), b.Y, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  gmk_mlp g{};
  unpack(G.data(), g);
  *grad = g;
  if (loss) {
    double l = 0;
    for (uint32_t s = 0; s < n; ++s) l += (double)(yh[s] - y[s]) * (yh[s] - y[s]);
    *loss = (float)(l / n);
  }
  return GMK_OK;
}

extern "C" int gmk_set_cost_model(gmk_ctx* c, const gmk_mlp* model) {
  if (!c) return GMK_EINVAL;
  c->has_model = model != nullptr;
  if (model) c->model = *model;
  return GMK_OK;
}

// Predicted cost of each graph: the model's prediction, or the size heuristic
// (edges x vertices) when no model is set.
int gmk_predict_costs(gmk_ctx* c, const gmk_graph_desc* graphs, uint32_t n, double* cost) {
  std::vector<float> X((size_t)n * IN);
  for (uint32_t i = 0; i < n; ++i) gmk_graph_features(&graphs[i], &X[(size_t)i * IN]);
  if (!c->has_model) {
    for (uint32_t i = 0; i < n; ++i) {
      const double e = graphs[i].off[graphs[i].n] / 2.0;
      cost[i] = e * (double)graphs[i].n;
    }
    return GMK_OK;
  }
  std::vector<float> y(n);
  int r = gmk_mlp_predict(c, &c->model, X.data(), n, y.data());
  if (r) return r;
  for (uint32_t i = 0; i < n; ++i) cost[i] = y[i];
  return GMK_OK;
}
