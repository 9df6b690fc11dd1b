// Host-side structures shared by the engine's translation units.
#pragma once
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
  bool keep = false;
  u64 limit = ~0ull;
  u64 res_cap = 0;
  double ms_expand = 0;
  u64 launches = 0;
  bool hit_limit = false;
  u64 fp_sum_extra = 0, fp_xor_extra = 0;
};

// Whole graph as one partition (local IDs = global IDs, right side shifted).
void build_global_partition(u32 n, u32 n_left, const u32* off, const u32* adj, u32 theta, int W, PartitionSet& ps);

// Appends a root state (header + 6W words) for partition pid.
u64* push_root(PartitionSet& ps, u32 pid);

}  // namespace gmk

int gmk_run_bb(gmk_ctx* c, const gmk::PartitionSet& ps, const gmk_params* prm, gmk_stats* st, gmk::RunState& rs);
int gmk_run_seeded(gmk_ctx* c, const gmk_params* prm, gmk_stats* st, gmk::RunState& rs);
