// Seed partitions for the results with both sides >= 2k+1 ("fat" results).
//
// For a seed s on side `us`, the partition holds every vertex that can be in, or
// be added to, a fat maximal k-biplex whose first vertex on side `us` (in the
// degeneracy order) is s.  Construction restates BPPivot's constructSets and
// pruneCX (baselines/MaximalBiPlex/src/biplex/biplex.cpp:227-393):
//   C_os = N(s);  C_us = later 2-hop vertices;  X_us = earlier 2-hop vertices;
//   degree pruning;  3-hop vertices of the other side with >= lb - k candidate
//   neighbours join C_os;  pruning again.
// Unlike the reference, a 3-hop vertex the seed misses joins C_os only when k >= 1:
// with k = 0 the seed cannot miss any vertex (the reference keeps them and loses
// results at k = 0).
#pragma once
#include <cstdint>
#include <vector>

namespace gmk {

struct HostGraph {
  uint32_t n, n_left;
  const uint32_t* off;
  const uint32_t* adj;
  uint32_t deg(uint32_t v) const { return off[v + 1] - off[v]; }
};

// Degeneracy order over all vertices (BPPivot coreOrdering, biplex.cpp:85-119):
// repeatedly remove a vertex of minimum remaining degree.  order[v] = position.
void degeneracy_order(const HostGraph& g, std::vector<uint32_t>& order);

struct SeedPartition {
  uint32_t seed = 0;
  int us = 0;                       // side of the seed (0 left, 1 right)
  std::vector<uint32_t> left, right;  // global IDs, ascending
  // root state membership, as flags parallel to left/right:
  //   0 = C, 1 = X, 2 = S (the seed)
  std::vector<uint8_t> lflag, rflag;
};

// Builds the partition of seed s; returns false when it cannot hold a result.
// scratch arrays must have n entries and be all-zero on entry (left zero on exit).
bool build_seed_partition(const HostGraph& g, const std::vector<uint32_t>& order, uint32_t s, uint32_t k,
                          uint32_t lb, SeedPartition& out, std::vector<int32_t>& cnt, std::vector<uint8_t>& mark);

}  // namespace gmk
