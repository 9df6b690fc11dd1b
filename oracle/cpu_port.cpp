// CPU port of the enumerator's thin-result method, multi-threaded with
// OpenMP: the multi-core CPU line of the benchmark (bench.py's cpu_baseline
// leg), next to the reference BPPivot on one core.  TEST AND BENCHMARK
// INFRASTRUCTURE ONLY: nothing in the product library links or calls it.
//
// It runs the same family construction, guard tables and maximality test as
// the HIP kernels (gmkbe_amd/csrc/thin.h, thin_build.h, compiled here for the
// host):
// one-vertex tuples of the left side in ID order (the heuristic order of the
// GPU's batched mode), each tuple's families in order, and every family's
// combinations in rank order split into chunks over the threads.  A run stops
// at the result limit (or when the one-vertex tuples are exhausted).  This is
// the part of the method that a limited run on the large power-law graphs
// reaches; the reference's own algorithm (BPPivot) searches the whole graph
// from one root at theta <= 2k and has no per-seed parallelism to port.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <vector>

#include "../gmkbe_amd/csrc/thin.h"
#include "../gmkbe_amd/csrc/thin_build.h"

using namespace gmk::thin;

namespace {
struct HostArena {
  std::vector<std::vector<u32>> blocks;
  u64 alloc(u64 n) { blocks.emplace_back(n ? n : 1); return blocks.size() - 1; }
  u32* ptr(u64 id) { return blocks[id].data(); }
};
struct HostFams {
  std::vector<Family> f;
  Family* next() { f.emplace_back(); return &f.back(); }
};
constexpr u64 CHUNK = 4096;  // ranks per work unit
constexpr u64 GROUP = 1024;  // work units per parallel loop
}  // namespace

// Returns the results found (at most `limit`, 0 = no limit) with `threads`
// threads (0: all), stopping after about `budget` seconds when budget > 0
// (whole chunks: a bounded sample); *seconds receives the enumeration time
// and *ranks the combinations tested.
extern "C" long long cpu_thin_port(unsigned n, unsigned n_left, const unsigned* off, const unsigned* adj, unsigned k,
                                   unsigned long long limit, int threads, double budget, double* seconds,
                                   unsigned long long* ranks) {
  if (k < 1 || k > 3) return -1;
  Graph g;
  g.off = off; g.adj = adj;
  g.a_lo = 0; g.a_hi = n_left; g.b_lo = n_left; g.b_hi = n;
  if (threads > 0) omp_set_num_threads(threads);
  const u64 lim = limit ? limit : ~0ull;
  std::atomic<u64> count{0}, tested{0};
  const auto t0 = std::chrono::steady_clock::now();
  std::atomic<bool> out_of_time{false};
  auto over = [&]() {
    if (budget <= 0) return false;
    if (out_of_time.load(std::memory_order_relaxed)) return true;
    if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > budget) out_of_time = true;
    return out_of_time.load();
  };
  for (u32 s = g.a_lo; s < g.a_hi && count.load() < lim && !over(); ++s) {
    const u32 L[1] = {s};
    HostArena ar;
    HostFams fo;
    Cost cost;
    build_families(g, L, 1, k, 1, 0, ar, fo, cost);
    for (Family& f : fo.f) {
      if (over()) break;
      guard_build(g, f, k, ar, cost);
    }
    for (const Family& f : fo.f) {
      if (count.load() >= lim || over()) break;
      const u64 nchunks = (f.total + CHUNK - 1) / CHUNK;
      // chunks in groups, so a family of 10^16 combinations stops at the
      // limit or the budget instead of skipping through every chunk
      for (u64 c0 = 0; c0 < nchunks && count.load() < lim && !over(); c0 += GROUP) {
      const u64 c1 = std::min(nchunks, c0 + GROUP);
#pragma omp parallel for schedule(dynamic, 1)
      for (u64 c = c0; c < c1; ++c) {
        if (count.load(std::memory_order_relaxed) >= lim || over()) continue;
        u64 found = 0, done = 0;
        Cost cc;
        const u64 r1 = std::min(f.total, (c + 1) * CHUNK);
        for (u64 r = c * CHUNK; r < r1; ++r) {
          u64 rr = r;
          u32 ch[18], ct[18], nch = 0;
          for (u32 t = 0; t < f.ntypes; ++t) {
            const u64 radix = binom(f.psize[t], f.cnt[t]);
            const u64 rt = rr % radix;
            rr /= radix;
            u32 id[3];
            unrank(rt, f.cnt[t], f.psize[t], id);
            for (u32 q = 0; q < f.cnt[t]; ++q) {
              ch[nch] = pool_member(f, t, id[q], g.b_lo, cc);
              ct[nch] = popc(f.mask[t]);
              ++nch;
            }
          }
          ++done;
          if (side_a_maximal<18>(g, fam_ref(f), ch, ct, nch, k, cc) == 1) ++found;
        }
        tested += done;
        count += found;
      }
      }
    }
  }
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (ranks) *ranks = tested.load();
  return (long long)std::min<u64>(count.load(), lim);
}
