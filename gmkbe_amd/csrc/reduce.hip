// (alpha, alpha)-core reduction of the loaded graph on the device, before the
// seed partitions of a seeded run are built (BPPivot's coreReduction,
// baselines/MaximalBiPlex/src/biplex/biplex.cpp:123-174, applied when lb > k
// with alpha = lb - k, biplex.cpp:441-450).  Every result with both sides
// >= lb, and every vertex that could be added to one, keeps at least lb - k
// neighbours inside the core, so the fat results of the core are those of the
// graph.
//
// Peeling runs frontier by frontier: a wave per removed vertex walks its list
// and decrements each neighbour's degree with one atomic; the decrement that
// takes a degree from alpha to alpha - 1 puts that neighbour on the next
// frontier (exactly once).  Then the surviving lists are compacted into a new
// CSR with the same vertex IDs (removed vertices keep empty lists): a count, a
// three-phase exclusive scan and an order-preserving ballot copy.
#include <algorithm>
#include <vector>

#include "gmk_host.h"

using namespace gmk;

namespace {

constexpr u32 SCAN_B = 1024;  // items per scan block (256 threads x 4)

__global__ __launch_bounds__(256) void core_init(const u32* off, u32 n, u32 thr, u32* deg, uint8_t* dead, u32* front,
                                                 u32* front_n) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool rm = false;
  if (v < n) {
    const u32 d = off[v + 1] - off[v];
    deg[v] = d;
    rm = d < thr;
    dead[v] = rm ? 1 : 0;
  }
  const u64 m = __ballot(rm);
  if (!m) return;
  const int leader = __ffsll((unsigned long long)m) - 1;
  u32 base = 0;
  if (lane == leader) base = atomicAdd(front_n, (u32)__popcll(m));
  base = (u32)__shfl((int)base, leader, 64);
  if (rm) front[base + (u32)__popcll(m & ((1ull << lane) - 1ull))] = v;
}

// One wave per vertex of the frontier.
__global__ __launch_bounds__(256) void core_peel(const u32* off, const u32* adj, const u32* front, u32 nf, u32 thr,
                                                 u32* deg, uint8_t* dead, u32* next, u32* next_n) {
  const u32 wv = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wv >= nf) return;
  const u32 v = front[wv];
  const u32 o = off[v], d = off[v + 1] - o;
  if (lane == 0) atomicAdd((unsigned long long*)(next_n + 1), (unsigned long long)d);  // entries walked
  for (u32 b = 0; b < d; b += 64) {
    bool push = false;
    u32 w = 0;
    if (b + lane < d) {
      w = adj[o + b + lane];
      if (!dead[w]) push = atomicSub(&deg[w], 1u) == thr;  // alpha -> alpha - 1: w leaves now
      if (push) dead[w] = 1;
    }
    const u64 m = __ballot(push);
    if (m) {
      const int leader = __ffsll((unsigned long long)m) - 1;
      u32 base = 0;
      if (lane == leader) base = atomicAdd(next_n, (u32)__popcll(m));
      base = (u32)__shfl((int)base, leader, 64);
      if (push) next[base + (u32)__popcll(m & ((1ull << lane) - 1ull))] = w;
    }
  }
}

// Exclusive scan of the surviving degrees (removed vertices: 0) into nofs;
// phase 1: per-block scans and block sums.
__global__ __launch_bounds__(256) void core_scan1(const u32* deg, const uint8_t* dead, u32 n, u32* nofs, u32* bsum) {
  __shared__ u32 ws[4];
  const u32 t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const u32 i0 = blockIdx.x * SCAN_B + t * 4;
  u32 x[4], s = 0;
  for (int q = 0; q < 4; ++q) {
    const u32 i = i0 + q;
    x[q] = (i < n && !dead[i]) ? deg[i] : 0u;
    s += x[q];
  }
  u32 inc = s;
  for (int d = 1; d < 64; d <<= 1) {
    const u32 o = __shfl_up(inc, d, 64);
    if ((int)lane >= d) inc += o;
  }
  if (lane == 63) ws[wv] = inc;
  __syncthreads();
  u32 before = 0;
  for (u32 q = 0; q < wv; ++q) before += ws[q];
  u32 run = before + inc - s;
  for (int q = 0; q < 4; ++q) {
    const u32 i = i0 + q;
    if (i < n) nofs[i] = run;
    run += x[q];
  }
  if (t == 255) bsum[blockIdx.x] = run;
}

// Phase 2: exclusive scan of the block sums (one block; the last entry gets
// the total).
__global__ __launch_bounds__(1024) void core_scan2(u32* bsum, u32 nb) {
  __shared__ u32 part[1024];
  const u32 t = threadIdx.x;
  const u32 seg = (nb + 1023) / 1024, b = t * seg < nb ? t * seg : nb, e = b + seg < nb ? b + seg : nb;
  u32 s = 0;
  for (u32 i = b; i < e; ++i) s += bsum[i];
  part[t] = s;
  __syncthreads();
  for (u32 d = 1; d < 1024; d <<= 1) {
    const u32 v = t >= d ? part[t] + part[t - d] : part[t];
    __syncthreads();
    part[t] = v;
    __syncthreads();
  }
  u32 run = t ? part[t - 1] : 0;
  for (u32 i = b; i < e; ++i) { const u32 v = bsum[i]; bsum[i] = run; run += v; }
  if (t == 1023) bsum[nb] = part[1023];
}

// Phase 3: block offsets added; the end offset written.
__global__ __launch_bounds__(256) void core_scan3(u32* nofs, u32 n, const u32* bsum) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) nofs[i] += bsum[i / SCAN_B];
  if (i == 0) nofs[n] = bsum[(n + SCAN_B - 1) / SCAN_B];
}

// One wave per surviving vertex: its surviving neighbours, in order.
__global__ __launch_bounds__(256) void core_copy(const u32* off, const u32* adj, u32 n, const uint8_t* dead,
                                                 const u32* nofs, u32* nadj) {
  const u32 v = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (v >= n || dead[v]) return;
  const u32 o = off[v], d = off[v + 1] - o;
  u32 at = nofs[v];
  for (u32 b = 0; b < d; b += 64) {
    bool keep = false;
    u32 w = 0;
    if (b + lane < d) { w = adj[o + b + lane]; keep = !dead[w]; }
    const u64 m = __ballot(keep);
    if (keep) nadj[at + (u32)__popcll(m & ((1ull << lane) - 1ull))] = w;
    at += (u32)__popcll(m);
  }
}

}  // namespace

int gmk::core_reduce(gmk_ctx* c, u32 thr, RunState& rs, CoreGraph& out) {
  const u32 n = c->n;
  const u64 m = c->adj.size();
  hipStream_t s = c->stream;
  int r;
  const u32 nb = (n + SCAN_B - 1) / SCAN_B;
  if ((r = ensure(c, c->core_deg, (size_t)n * 4 + 32))) return r;
  if ((r = ensure(c, c->core_dead, (size_t)n + 16))) return r;
  if ((r = ensure(c, c->core_front, 2ull * n * 4 + 16))) return r;
  if ((r = ensure(c, c->core_off, (size_t)(n + 1) * 4 + (size_t)(nb + 1) * 4 + 16))) return r;
  u32* deg = (u32*)c->core_deg.p;
  u32* d_nf = deg + n + (n & 1);  // [0] first frontier, [1] next frontier, [2..3] entries walked (u64); 8-aligned
  uint8_t* dead = (uint8_t*)c->core_dead.p;
  u32* fr[2] = {(u32*)c->core_front.p, (u32*)c->core_front.p + n};
  u32* nofs = (u32*)c->core_off.p;
  u32* bsum = nofs + n + 1;
  HIPCHK(c, hipEventRecord(c->ev2, s));
  HIPCHK(c, hipMemsetAsync(d_nf, 0, 16, s));
  const dim3 b(256);
  hipLaunchKernelGGL(core_init, dim3((n + 255) / 256), b, 0, s, c->d_off, n, thr, deg, dead, fr[0], d_nf);
  HIPCHK(c, hipGetLastError());
  u32 nf = 0;
  HIPCHK(c, hipMemcpyAsync(&nf, d_nf, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  u32 removed = nf, rounds = 0;
  int cur = 0;
  while (nf) {
    HIPCHK(c, hipMemsetAsync(d_nf + 1, 0, 4, s));
    hipLaunchKernelGGL(core_peel, dim3((nf + 3) / 4), b, 0, s, c->d_off, c->d_adj, fr[cur], nf, thr, deg, dead,
                       fr[cur ^ 1], d_nf + 1);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&nf, d_nf + 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    cur ^= 1;
    removed += nf;
    ++rounds;
  }
  // compaction
  hipLaunchKernelGGL(core_scan1, dim3(nb), b, 0, s, deg, dead, n, nofs, bsum);
  hipLaunchKernelGGL(core_scan2, dim3(1), dim3(1024), 0, s, bsum, nb);
  hipLaunchKernelGGL(core_scan3, dim3((n + 255) / 256), b, 0, s, nofs, n, bsum);
  HIPCHK(c, hipGetLastError());
  u32 m2 = 0;
  u64 walked = 0;
  HIPCHK(c, hipMemcpyAsync(&m2, nofs + n, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(&walked, d_nf + 2, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if ((r = ensure(c, c->core_adj, (size_t)std::max<u32>(m2, 1) * 4))) return r;
  hipLaunchKernelGGL(core_copy, dim3((n + 3) / 4), b, 0, s, c->d_off, c->d_adj, n, dead, nofs, (u32*)c->core_adj.p);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev3, s));
  out.off.resize((size_t)n + 1);
  out.adj.resize(m2);
  HIPCHK(c, hipMemcpyAsync(out.off.data(), nofs, (size_t)(n + 1) * 4, hipMemcpyDeviceToHost, s));
  if (m2) HIPCHK(c, hipMemcpyAsync(out.adj.data(), c->core_adj.p, (size_t)m2 * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev2, c->ev3));
  // Algorithmic bytes.  Init: offsets read, degree and flag written (9 B per
  // vertex) and 4 B per removed vertex queued.  Peel: per removed vertex its
  // offsets and queue entry (12 B) and its list, per entry 4 B read, the flag
  // (1 B) and the degree atomic (8 B).  Scan: degree and flag read, offset
  // written and updated (17 B per vertex).  Copy: offsets of the surviving
  // vertices, their lists with every entry's flag (5 B per entry) and the
  // surviving entries written.
  const u64 live_entries = m - walked;  // entries of surviving vertices' lists
  const u64 bytes = (u64)n * 9 + (u64)removed * 4 + (u64)removed * 12 + walked * 13 + (u64)n * 17 +
                    (u64)(n - removed) * 8 + live_entries * 5 + (u64)m2 * 4;
  out.d_off = nofs;
  out.d_adj = (const u32*)c->core_adj.p;
  out.removed = removed;
  rs.ms_reduce += ms;
  rs.bytes_reduce += bytes;
  rs.reduce_removed += removed;
  rs.reduce_rounds += rounds;
  return GMK_OK;
}
