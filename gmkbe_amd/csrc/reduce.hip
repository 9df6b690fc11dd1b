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

// ---- Butterfly reduction (BPPivot butterflyReduction, biplex.cpp:176-225,
// applied when lb > 2k, biplex.cpp:469-478).  For a vertex u, cn[w] counts the
// common neighbours of u and each w on u's side; a neighbour v supports u when
// at least q - k of v's neighbours w have cn[w] >= q - 2k.  An unsupported
// edge (u, v) is in no k-biplex with both sides >= q, and neither is a vertex
// with fewer than q - k supporting neighbours: in such a biplex H, v misses at
// most k of the members on u's side, and each of them shares at least q - 2k
// neighbours in H with u.  The rule keeps every large maximal result and its
// maximality (every edge and vertex of a larger biplex extending it stays).
// BPPivot takes the vertices one at a time; here every vertex of a side is
// decided in parallel against the removals of the previous pass (counts on a
// larger graph are never smaller, so every decision stays sound), with
// BPPivot's two rounds over both sides.  One workgroup per vertex: waves take
// u's neighbours, lanes their lists; the counters are the workgroup's own
// slice of an n-entry array, cleared by walking the same lists again.
struct BflyArgs {
  const u32* off;
  const u32* adj;
  u32 n, lo, hi;            // all vertices; this pass's side
  u32 need_cn, need_sup;    // q - 2k, q - k
  const uint8_t* rm_old;    // removed before this pass
  uint8_t* rm_new;          // removed after it (a copy of rm_old on entry)
  uint8_t* emask;           // per list entry: 1 when its edge is removed
  u32* deg;                 // degrees, less the removed neighbours
  u32* cn;                  // gridDim.x slices of n counters (zero on entry and exit)
  u32* changed;             // [0] vertices, [1] edges removed in this pass
  unsigned long long* bytes;
};

__global__ __launch_bounds__(256) void bfly_pass(BflyArgs a) {
  __shared__ u32 support;
  u32* cn = a.cn + (u64)blockIdx.x * a.n;
  const u32 wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  u64 bytes = 0;
  for (u32 u = a.lo + blockIdx.x; u < a.hi; u += gridDim.x) {
    if (a.rm_old[u]) continue;
    const u32 o = a.off[u], d = a.off[u + 1] - o;
    if (threadIdx.x == 0) support = 0;
    for (u32 i = wv; i < d; i += 4) {
      const u32 v = a.adj[o + i];
      if (a.rm_old[v]) continue;
      const u32 ov = a.off[v], dv = a.off[v + 1] - ov;
      for (u32 j = lane; j < dv; j += 64) {
        const u32 w = a.adj[ov + j];
        if (!a.rm_old[w]) atomicAdd(&cn[w], 1u);
      }
      bytes += 9ull * dv;
    }
    __syncthreads();
    for (u32 i = wv; i < d; i += 4) {
      const u32 v = a.adj[o + i];
      if (a.rm_old[v]) continue;
      const u32 ov = a.off[v], dv = a.off[v + 1] - ov;
      u32 c = 0;
      for (u32 j = lane; j < dv; j += 64) {
        const u32 w = a.adj[ov + j];
        if (!a.rm_old[w] && cn[w] >= a.need_cn) ++c;
      }
      c = (u32)wave_sum_u64(c);
      bytes += 9ull * dv;
      if (lane == 0) {
        if (c >= a.need_sup) {
          atomicAdd(&support, 1u);
        } else {
          // the edge leaves: its entry in N(u) and, found by binary search, in N(v)
          a.emask[o + i] = 1;
          u32 lo = 0, hi = dv;
          while (lo < hi) {
            const u32 mid = (lo + hi) >> 1;
            if (a.adj[ov + mid] < u) lo = mid + 1; else hi = mid;
          }
          if (lo < dv && a.adj[ov + lo] == u) a.emask[ov + lo] = 1;
          atomicAdd(&a.changed[1], 1u);
        }
      }
    }
    __syncthreads();
    if (support < a.need_sup) {
      // u leaves; a neighbour whose degree falls below q - k leaves with it
      if (threadIdx.x == 0) {
        a.rm_new[u] = 1;
        atomicAdd(&a.changed[0], 1u);
      }
      for (u32 i = threadIdx.x; i < d; i += blockDim.x) {
        const u32 v = a.adj[o + i];
        if (atomicSub(&a.deg[v], 1u) == a.need_sup) {
          a.rm_new[v] = 1;
          atomicAdd(&a.changed[0], 1u);
        }
      }
    }
    for (u32 i = wv; i < d; i += 4) {
      const u32 v = a.adj[o + i];
      if (a.rm_old[v]) continue;
      const u32 ov = a.off[v], dv = a.off[v + 1] - ov;
      for (u32 j = lane; j < dv; j += 64) cn[a.adj[ov + j]] = 0;
    }
    __syncthreads();
  }
  const u64 b = wave_sum_u64(bytes);
  if (lane == 0 && b) atomicAdd(a.bytes, (unsigned long long)b);
}

// Degrees of the input graph; vertices without neighbours (those the core
// reduction removed, and isolated ones) start removed.
__global__ __launch_bounds__(256) void bfly_init(const u32* off, u32 n, u32* deg, uint8_t* rm) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const u32 d = off[v + 1] - off[v];
  deg[v] = d;
  rm[v] = d == 0 ? 1 : 0;
}

// Surviving degree of every vertex (entries whose edge and far end remain),
// and the totals: vertices removed by the passes ([0]) and removed entries
// between surviving vertices ([1], two per edge).
__global__ __launch_bounds__(256) void bfly_count(const u32* off, const u32* adj, u32 n, const uint8_t* rm,
                                                  const uint8_t* emask, u32* deg, u32* tot) {
  const u32 v = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (v >= n) return;
  const u32 o = off[v], d = off[v + 1] - o;
  u32 c = 0, cut = 0;
  if (!rm[v])
    for (u32 j = lane; j < d; j += 64) {
      const bool far = !rm[adj[o + j]];
      c += (far && !emask[o + j]) ? 1u : 0u;
      cut += (far && emask[o + j]) ? 1u : 0u;
    }
  c = (u32)wave_sum_u64(c);
  cut = (u32)wave_sum_u64(cut);
  if (lane == 0) {
    deg[v] = c;
    if (rm[v] && d > 0) atomicAdd(&tot[0], 1u);
    if (cut) atomicAdd(&tot[1], cut);
  }
}

// One wave per surviving vertex: its surviving entries, in order.
__global__ __launch_bounds__(256) void bfly_copy(const u32* off, const u32* adj, u32 n, const uint8_t* rm,
                                                 const uint8_t* emask, const u32* nofs, u32* nadj) {
  const u32 v = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (v >= n || rm[v]) return;
  const u32 o = off[v], d = off[v + 1] - o;
  u32 at = nofs[v];
  for (u32 b = 0; b < d; b += 64) {
    bool keep = false;
    u32 w = 0;
    if (b + lane < d) { w = adj[o + b + lane]; keep = !emask[o + b + lane] && !rm[w]; }
    const u64 m = __ballot(keep);
    if (keep) nadj[at + (u32)__popcll(m & ((1ull << lane) - 1ull))] = w;
    at += (u32)__popcll(m);
  }
}

}  // namespace

int gmk::butterfly_reduce(gmk_ctx* c, const CoreGraph& in, u32 q, u32 k, RunState& rs, CoreGraph& out) {
  const u32 n = c->n;
  const u64 m = in.adj.size();
  hipStream_t s = c->stream;
  int r;
  const u32 nb = (n + SCAN_B - 1) / SCAN_B;
  // counter slices: at most 256 MiB of them
  const u32 grid = (u32)std::max<u64>(1, std::min<u64>(2ull * c->n_cu, (256ull << 20) / (4ull * std::max<u32>(n, 1))));
  if ((r = ensure(c, c->bfly_cnt, (size_t)grid * n * 4 + 16))) return r;
  if ((r = ensure(c, c->bfly_mark, 2ull * n + m + 64))) return r;
  if ((r = ensure(c, c->bfly_deg, (size_t)n * 4 + 64))) return r;
  if ((r = ensure(c, c->bfly_off, (size_t)(n + 1) * 4 + (size_t)(nb + 1) * 4 + 16))) return r;
  u32* cn = (u32*)c->bfly_cnt.p;
  uint8_t* rm[2] = {(uint8_t*)c->bfly_mark.p, (uint8_t*)c->bfly_mark.p + n};
  uint8_t* emask = (uint8_t*)c->bfly_mark.p + 2ull * n;
  u32* deg = (u32*)c->bfly_deg.p;
  u32* changed = deg + n + (n & 1);  // [0] vertices, [1] edges removed, [2..3] bytes walked (u64), [4..5] totals; 8-aligned
  u32* nofs = (u32*)c->bfly_off.p;
  u32* bsum = nofs + n + 1;
  HIPCHK(c, hipEventRecord(c->ev2, s));
  HIPCHK(c, hipMemsetAsync(cn, 0, (size_t)grid * n * 4, s));
  HIPCHK(c, hipMemsetAsync(c->bfly_mark.p, 0, 2ull * n + m, s));
  HIPCHK(c, hipMemsetAsync(changed, 0, 24, s));
  hipLaunchKernelGGL(bfly_init, dim3((n + 255) / 256), dim3(256), 0, s, in.d_off, n, deg, rm[0]);
  HIPCHK(c, hipGetLastError());
  u32 passes = 0;
  for (int t = 0; t < 2; ++t) {
    bool any = false;
    for (int side = 0; side < 2; ++side) {
      HIPCHK(c, hipMemsetAsync(changed, 0, 8, s));
      HIPCHK(c, hipMemcpyAsync(rm[1], rm[0], n, hipMemcpyDeviceToDevice, s));
      BflyArgs a{in.d_off, in.d_adj, n, side ? c->n_left : 0u, side ? n : c->n_left, q - 2 * k, q - k,
                 rm[0], rm[1], emask, deg, cn, changed, (unsigned long long*)(changed + 2)};
      hipLaunchKernelGGL(bfly_pass, dim3(grid), dim3(256), 0, s, a);
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipMemcpyAsync(rm[0], rm[1], n, hipMemcpyDeviceToDevice, s));
      u32 h[2] = {0, 0};
      HIPCHK(c, hipMemcpyAsync(h, changed, 8, hipMemcpyDeviceToHost, s));
      HIPCHK(c, hipStreamSynchronize(s));
      ++passes;
      any = any || h[0] || h[1];
    }
    if (!any) break;
  }
  // compaction (removed vertices keep empty lists)
  hipLaunchKernelGGL(bfly_count, dim3((n + 3) / 4), dim3(256), 0, s, in.d_off, in.d_adj, n, rm[0], emask, deg,
                     changed + 4);
  hipLaunchKernelGGL(core_scan1, dim3(nb), dim3(256), 0, s, deg, rm[0], n, nofs, bsum);
  hipLaunchKernelGGL(core_scan2, dim3(1), dim3(1024), 0, s, bsum, nb);
  hipLaunchKernelGGL(core_scan3, dim3((n + 255) / 256), dim3(256), 0, s, nofs, n, bsum);
  HIPCHK(c, hipGetLastError());
  u32 m2 = 0, tot[2] = {0, 0};
  u64 walked = 0;
  HIPCHK(c, hipMemcpyAsync(&m2, nofs + n, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(&walked, changed + 2, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(tot, changed + 4, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  const u32 removed = tot[0], edges = tot[1] / 2;
  if ((r = ensure(c, c->bfly_adj, (size_t)std::max<u32>(m2, 1) * 4))) return r;
  hipLaunchKernelGGL(bfly_copy, dim3((n + 3) / 4), dim3(256), 0, s, in.d_off, in.d_adj, n, rm[0], emask, nofs,
                     (u32*)c->bfly_adj.p);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(c->ev3, s));
  out.off.resize((size_t)n + 1);
  out.adj.resize(m2);
  HIPCHK(c, hipMemcpyAsync(out.off.data(), nofs, (size_t)(n + 1) * 4, hipMemcpyDeviceToHost, s));
  if (m2) HIPCHK(c, hipMemcpyAsync(out.adj.data(), c->bfly_adj.p, (size_t)m2 * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev2, c->ev3));
  // Algorithmic bytes: the passes' list walks (9 B per entry: the entry, the
  // far end's flag and its counter) as counted in-kernel; the initial degrees
  // (9 B per vertex); per count or copy round the offsets and flags (9 B per
  // vertex) and every entry with its mask and far flag (6 B); the scan (17 B
  // per vertex) and the surviving entries written.
  const u64 bytes = walked + (u64)n * 9 + 2 * ((u64)n * 9 + m * 6) + (u64)n * 17 + (u64)m2 * 4;
  out.d_off = nofs;
  out.d_adj = (const u32*)c->bfly_adj.p;
  out.removed = in.removed + removed;
  rs.ms_reduce += ms;
  rs.bytes_reduce += bytes;
  rs.reduce_removed += removed;
  rs.reduce_edges += edges;
  rs.reduce_rounds += passes;
  return GMK_OK;
}

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
