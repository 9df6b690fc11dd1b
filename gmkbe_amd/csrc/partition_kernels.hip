// Bitset adjacency rows of seed partitions, built on the device.
//
// A seed partition (BPPivot constructSets, biplex.cpp:315-393) is a local
// universe: nl left and nr right vertices, each side listed by ascending global
// ID in the map pool.  Its rows are, for every local vertex, the bitset of its
// neighbours on the other side in local IDs (nl rows of wr words, then nr rows
// of wl words).  The host computes the vertex sets; this kernel fills the rows
// from the graph's CSR: one workgroup of four waves per partition, one wave per
// row at a time; lanes split the row's adjacency list, find each neighbour's
// local ID by binary search in the other side's map, and set its bit in an LDS
// copy of the row, which the wave then writes out with coalesced stores.  Rows
// of up to 64 words use 2 KB of LDS per workgroup; a launch with wider
// partitions (up to 1024 words) uses the 32 KB form.
#include "gmk_device.h"

namespace gmk {

struct RowBuildArgs {
  const PartDesc* parts;
  const u32* pids;      // partitions to build
  u32 n;
  const u32* maps;
  const u32* off;       // CSR offsets of all graphs, concatenated
  const u32* adj;       // CSR lists of all graphs, concatenated
  const u64* off_base;  // per graph: start of its offsets in off
  const u64* adj_base;  // per graph: start of its lists in adj
  u64* rows;
  u64* bytes;           // algorithmic bytes (adjacency + probes + rows)
  u32 max_words;        // widest row of the partitions built
};

template <int MAXW>
__global__ __launch_bounds__(256) void build_part_rows(RowBuildArgs a) {
  __shared__ u64 row[4][MAXW];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (blockIdx.x >= a.n) return;
  const PartDesc P = a.parts[a.pids[blockIdx.x]];
  const u32* off = a.off + a.off_base[P.graph];
  const u32* adj = a.adj + a.adj_base[P.graph];
  u64 bytes = 0;
  const u32 rows_total = P.nl + P.nr;
  for (u32 r = wv; r < rows_total; r += 4) {
    const bool left = r < P.nl;
    const u32 i = left ? r : r - P.nl;
    const u32 v = a.maps[(left ? P.map_l : P.map_r) + i];
    const u32* other = a.maps + (left ? P.map_r : P.map_l);
    const u32 n_other = left ? P.nr : P.nl;
    const u32 w_row = left ? P.wr : P.wl;
    u64* out = a.rows + (left ? P.rows_l + (u64)i * P.wr : P.rows_r + (u64)i * P.wl);
    for (u32 w = lane; w < w_row; w += 64) row[wv][w] = 0;
    __builtin_amdgcn_wave_barrier();
    const u32 e0 = off[v], e1 = off[v + 1];
    for (u32 e = e0 + lane; e < e1; e += 64) {
      const u32 x = adj[e];
      u32 lo = 0, hi = n_other;
      while (lo < hi) {
        const u32 mid = (lo + hi) >> 1;
        if (other[mid] < x) lo = mid + 1; else hi = mid;
        bytes += 4;
      }
      if (lo < n_other && other[lo] == x) atomicOr((unsigned long long*)&row[wv][lo >> 6], 1ull << (lo & 63));
    }
    bytes += 4ull * (e1 - e0 > (u32)lane ? (e1 - e0 - lane + 63) / 64 : 0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (u32 w = lane; w < w_row; w += 64) {
      out[w] = row[wv][w];
      bytes += 8;
    }
    __builtin_amdgcn_wave_barrier();
  }
  const u64 b = wave_sum_u64(bytes);
  if (lane == 0 && b) atomicAdd((unsigned long long*)a.bytes, (unsigned long long)b);
}

hipError_t launch_build_part_rows(const RowBuildArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.max_words > 1024) return hipErrorInvalidValue;
  if (a.max_words <= 64) hipLaunchKernelGGL(build_part_rows<64>, dim3(a.n), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(build_part_rows<1024>, dim3(a.n), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace gmk
