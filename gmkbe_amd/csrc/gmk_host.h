// Host-side structures shared by the engine's translation units.
#pragma once
#include <string>
#include <vector>

#include "gmk_device.h"
#include "../../include/gmk.h"

namespace gmk {

// Partitions of one run, with their root search states.
struct PartitionSet {
  int W = 0;                     // bitset words per set (1, 2, 4, 8 or 16)
  std::vector<PartDesc> parts;
  std::vector<u64> rows;         // row pool (bitset adjacency)
  std::vector<u32> maps;         // local -> global ID maps
  std::vector<u64> roots;        // root states, 1 + 6W words each
};

struct RunState {
  DedupTable dedup{};
  unsigned long long* gcount = nullptr;  // per-graph caps (multi-graph batches)
  u64* gfp = nullptr;
  u64 gcap = ~0ull;
  bool keep = false;
  u64 limit = ~0ull;
  u64 res_cap = 0;
  double ms_expand = 0;
  u64 launches = 0;
  bool hit_limit = false;
  u64 fp_sum_extra = 0, fp_xor_extra = 0;
  double ms_kernel[4] = {0, 0, 0, 0};
  u64 launches_kernel[4] = {0, 0, 0, 0};
  // progress log: results so far and elapsed ms after each launch
  std::vector<u64> prog_count;
  std::vector<double> prog_ms;
  double t0_ms = 0;
  void progress(u64 count);
};

// Whole graph as one partition (local IDs = global IDs, right side shifted).
void build_global_partition(u32 n, u32 n_left, const u32* off, const u32* adj, u32 k, u32 theta, int W, PartitionSet& ps,
                            u32 graph = 0);

// Appends a root state (header + 6W words) for partition pid.
u64* push_root(PartitionSet& ps, u32 pid);

}  // namespace gmk

using gmk::u32;
using gmk::u64;
using gmk::NSTRIPE;

#define HIPCHK(ctx, call)                                                                    \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                       \
      return GMK_EDEVICE;                                                                   \
    }                                                                                       \
  } while (0)

// Device counters of one run, read back after every batch.
struct RunCounters {
  u32 out_n, retry_n, flags, pad;  // pad: work counter of the persistent expansion kernel
  u64 res_words, res_count;
  u64 fp[NSTRIPE * 2];
  u64 stat[NSTRIPE * 8];  // 0 states, 1 intersections, 2 emitted leaves, 3-6 bytes per kernel
  u64 dstat[NSTRIPE * 4]; // dedup: inserts, probes, duplicates
};

struct DeviceVec {
  void* p = nullptr;
  size_t bytes = 0;
};

struct gmk_ctx {
  int device = 0;
  u32 persistent_waves = 256 * 16;  // waves of the persistent expansion kernel (16 per CU)
  hipStream_t stream = nullptr;
  std::string err;
  // graph (host copy kept for partition construction)
  u32 n = 0, n_left = 0;
  std::vector<u32> off, adj;
  u32* d_off = nullptr;
  u32* d_adj = nullptr;
  // per-run buffers, grown on demand and reused
  DeviceVec stack, next, retry, rows, maps, parts, res, counters;
  DeviceVec tuples, arena, fams, prefix, tcnt;  // thin-result kernels
  DeviceVec dedup;                              // result fingerprint table
  DeviceVec gcnt;                               // per-graph counters of a batch
  std::vector<u32> order;                        // degeneracy order (seeded runs)
  bool order_valid = false;
  RunCounters* h_cnt = nullptr;  // pinned
  u64 res_words = 0;             // record stream length of the last run
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<u64> prog_count;   // progress log of the last run
  std::vector<double> prog_ms;
  bool has_model = false;        // cost model for GMK_MODE_LEARNED
  gmk_mlp model{};
};

inline int ensure(gmk_ctx* c, DeviceVec& v, size_t bytes) {
  if (v.bytes >= bytes) return GMK_OK;
  if (v.p) (void)hipFree(v.p);
  v.p = nullptr; v.bytes = 0;
  size_t b = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&v.p, b);
  if (e != hipSuccess) {
    c->err = "hipMalloc(" + std::to_string(b) + " B) failed: " + hipGetErrorString(e);
    return GMK_ENOMEM;
  }
  v.bytes = b;
  return GMK_OK;
}


int gmk_pick_w(uint32_t words);
int gmk_predict_costs(gmk_ctx* c, const gmk_graph_desc* graphs, uint32_t n, double* cost);
int gmk_run_bb(gmk_ctx* c, const gmk::PartitionSet& ps, const gmk_params* prm, gmk_stats* st, gmk::RunState& rs);
int gmk_run_seeded(gmk_ctx* c, const gmk_params* prm, gmk_stats* st, gmk::RunState& rs);
// Thin results with the tuple on side_a (0 left, 1 right); generated side >= min_b.
int gmk_run_thin(gmk_ctx* c, const gmk_params* prm, gmk_stats* st, gmk::RunState& rs, int side_a, uint32_t min_b);
