// Kernel arguments of the thin-result kernels (shared by host and device code).
#pragma once
#include "gmk_device.h"
#include "thin.h"

namespace gmk {

struct ThinArgs {
  thin::Graph g;
  u32 k;
  u32 min_b;            // minimum size of the generated side
  // T1
  const u32* tuples;    // ntuples x j
  u32 ntuples, j;
  u32* arena;
  u64 arena_cap;
  u64* arena_used;
  thin::Family* fams;
  u32 fam_cap;
  u32* fam_n;
  u32* tflags;          // family-construction overflow flags
  u32* work;            // chunk counter of a check launch (zeroed before it)
  u32 budget;           // candidates a lane tests before deferring the rest of a rank
  u32 wave1_max;        // one-vertex tuples per launch below which a wave builds each tuple
  u32* tcount;          // that build: families per tuple, their offsets, 6 words kept per tuple (8 x tuples)
  u64* defer;           // deferred ranks of a check launch (capacity: the launch's ranks)
  u32* defer_n;         // their number (zeroed before the launch)
  u32* dwork;           // next deferred rank to take (zeroed before the launch)
  u32* flags;           // run flags (limit, result overflow)
  // T2
  const u64* prefix;    // nfam + 1 prefix sums of family totals
  const u32* ridx;      // per 2^RIDX_LOG ranks of the launch's range: the family of its first rank
  u32 nfam;
  u64 r0, r1;           // rank range of this launch
  int unlimited;        // 1 when the launch cannot reach the result limit (no slot check)
  u32 guard_lists;      // 1: ranks of families with guard tables may take the guard lists
  u64 guard_min_ranks;  // families with fewer combinations get no guard table
  u64 rank_base;        // offset of the launch's ranks inside every family (rank slices)
  int keep;
  int swap_sides;       // 1 when side A is the right side (records list L first)
  u32* res;
  u64 res_cap;
  u64* res_words;
  u64* res_count;
  u64 limit;
  u64* fp;
  u64* stat;
  DedupTable dedup;
  // Item log (probe runs of the learned schedule, null otherwise): per tuple of
  // the launch, results emitted and device lane time (wall-clock ticks).
  u64* item_log;
};

hipError_t launch_thin_build(const ThinArgs& a, hipStream_t s);
// Launch forms of the thin check (gmk_tuning.check_form): one lane per rank over
// the whole grid, or waves of a fixed grid taking 64-rank chunks.
enum : u32 { CHECK_GRID = 0, CHECK_CHUNKED = 1 };
hipError_t launch_thin_check(const ThinArgs& a, hipStream_t s, u32 max_blocks, u32 form);
hipError_t launch_thin_check_deferred(const ThinArgs& a, hipStream_t s, u32 blocks);
// Guard tables of the families just built (their number read from a.fam_n on
// the device; `blocks` workgroups of 4 waves take them in turn); runs before the
// family prefix sums, since a dominated family gets total 0.
hipError_t launch_thin_guard(const ThinArgs& a, hipStream_t s, u32 blocks);
// Neighbour hashes from nv work items (vertex, first neighbour index), each
// inserting up to NH_CHUNK neighbours of its vertex (slots already NH_EMPTY).
constexpr u32 NH_CHUNK = 2048;
hipError_t launch_build_neighbour_hash(const u32* off, const u32* adj, const u32* items, u32 nv, const u64* nh_off,
                                       u32* nh, hipStream_t s);
// Launches of fewer one-vertex tuples than this build them a wave per tuple (a
// lane per tuple would not fill the device); ThinArgs.wave1_max holds it.
constexpr u32 WAVE1_MAX_TUPLES = 16384;
// Prefix sums of the family totals over the rank slice [base, base + cap) of
// every family: min(max(total - base, 0), cap) (probe runs check a family's
// first ranks, learned runs take the families slice by slice; base 0, cap ~0:
// whole families).
hipError_t launch_thin_scan(const thin::Family* fams, u32 nfam, u64* prefix, u64 base, u64 cap, hipStream_t s);
// The same with the family count read on the device (*nfam_p, at most fam_cap);
// the sum goes to *total_out.
hipError_t launch_thin_scan_dev(const thin::Family* fams, const u32* nfam_p, u32 fam_cap, u64* prefix, u64 base,
                                u64 cap, u64* total_out, hipStream_t s);
// Rank index of a check launch: ridx[b] = family of rank min(r0 + b 2^RIDX_LOG,
// r1 - 1), b = 0 .. ceil((r1 - r0) / 2^RIDX_LOG).  It also clears the three
// consecutive words at zero3 (the check kernels' chunk counter, deferred-rank
// count and next deferred rank), in place of a memset launch.
constexpr u32 RIDX_LOG = 8;
hipError_t launch_thin_rank_index(const u64* prefix, u32 nfam, u64 r0, u64 r1, u32* ridx, u32* zero3,
                                  hipStream_t s);

}  // namespace gmk
