// Shared device-side definitions of the enumerator's HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gmk {

typedef uint32_t u32;
typedef uint64_t u64;

constexpr u32 NONE = 0xFFFFFFFFu;
constexpr u32 INVALID_PID = 0xFFFFFFFFu;
constexpr u32 FLAG_LIMIT = 1u;
constexpr u32 FLAG_RES_OVERFLOW = 2u;
constexpr u32 FLAG_TABLE_FULL = 4u;
constexpr int NSTRIPE = 64;  // striped counters (stat and fingerprint atomics)
constexpr int NSTAT = 8;     // counters per stripe

// Fingerprint constants: fp(L, R) = fmix64(sum mix64(id) + |L| A + |R| B).
constexpr u64 FP_A = 0x2545F4914F6CDD1Dull;
constexpr u64 FP_B = 0x9E6C63D0676A9A99ull;

__host__ __device__ __forceinline__ u64 mix64(u64 x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__host__ __device__ __forceinline__ u64 fmix64(u64 h) {
  h ^= h >> 33; h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33; return h;
}

// A partition: a local universe of nl left and nr right vertices (local IDs
// ordered like their global IDs) with bitset adjacency rows in the row pool.
struct PartDesc {
  u32 nl, nr;   // local universe sizes
  u32 wl, wr;   // 64-bit words per left / right bitset
  u64 rows_l;   // word offset in the row pool of the nl left rows (wr words each)
  u64 rows_r;   // word offset of the nr right rows (wl words each)
  u64 map_l;    // u32 offset in the map pool: local -> global, left
  u64 map_r;    // same, right
  u32 lb;       // both result sides must reach lb
  u32 seed;     // seed vertex (global ID) or NONE for a global partition
  u32 k;        // misses allowed per vertex
  u32 graph;    // graph of a multi-graph batch (0 otherwise)
};

// Second, independent per-vertex mix for the dedup table's check word.
__host__ __device__ __forceinline__ u64 mix64b(u64 x) { return fmix64(x * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull); }

// Result dedup table: open addressing over two independent 64-bit hashes of a
// result, kept in one 16-byte slot: word 0 the fingerprint (0 = empty slot),
// word 1 a check hash.  Linear probing from a slot taken from a second mix of
// the fingerprint; a slot is claimed by compare-and-swap on word 0 and its check
// word is written right after.  A result is a duplicate only when both words
// match, so two distinct results whose fingerprints collide are both kept.
struct DedupTable {
  unsigned long long* slots;  // 2 (mask + 1) words; capacity mask + 1, a power of two
  u64 mask;
  u64* stats;                 // [NSTRIPE][4]: inserts, probes, duplicates, undecided
  int on;
};
constexpr u32 DEDUP_SPIN = 1u << 14;  // bounded wait for a just-claimed slot's check word

// Returns false when (fp, fp2) is already in the table (the result is a
// duplicate).  A slot with the same fingerprint whose check word is not yet
// visible after DEDUP_SPIN polls leaves the result undecided: it is kept
// (never dropped on a guess) and counted in stats[3].
__device__ __forceinline__ bool dedup_insert(const DedupTable& t, u64 fp, u64 fp2, u32* flags, u32 stripe) {
  if (!t.on) return true;
  const unsigned long long key = fp ? fp : 1ull, key2 = fp2 ? fp2 : 1ull;
  u64 h = mix64(key ^ 0xD6E8FEB86659FD93ull) & t.mask;
  u64 probes = 0;
  int state = 0;  // 1 new, 2 duplicate, 3 undecided, 4 table full
  while (state == 0) {
    ++probes;
    unsigned long long* sl = t.slots + 2 * h;
    const unsigned long long old = atomicCAS(&sl[0], 0ull, key);
    if (old == 0ull) {
      __hip_atomic_store(&sl[1], key2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      state = 1;
    }
    if (state == 0 && old == key) {
      unsigned long long k2 = 0;
      for (u32 spin = 0; spin < DEDUP_SPIN; ++spin) {
        k2 = __hip_atomic_load(&sl[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (k2) break;
        __builtin_amdgcn_s_sleep(1);
      }
      if (k2 == key2) state = 2;
      else if (k2 == 0) state = 3;
    }
    if (state == 0) {
      h = (h + 1) & t.mask;
      if (probes > t.mask) state = 4;
    }
  }
  if (state == 4) atomicOr(flags, 4u /* FLAG_TABLE_FULL */);
  u64* st = t.stats + (stripe % NSTRIPE) * 4;
  atomicAdd((unsigned long long*)&st[0], 1ull);
  atomicAdd((unsigned long long*)&st[1], (unsigned long long)probes);
  if (state == 2) atomicAdd((unsigned long long*)&st[2], 1ull);
  if (state == 3) atomicAdd((unsigned long long*)&st[3], 1ull);
  return state != 2;
}

// Batch control of one state stack (one bitset width) in device memory, for
// steps planned on the device (bb_plan): the planned batch is the top n_in
// states, [in_off, in_off + n_in), and its children may take out_cap slots.
struct LaneCtl {
  u64 top;        // states on the stack
  u64 cap;        // its capacity in states
  u64 in_off;     // first state of the planned batch
  u32 n_in;       // states in the planned batch (0: the width is idle)
  u32 out_cap;    // children the batch may write
  u32 bmax;       // states per batch at most
  u32 halt;       // 0 running, 1 the stack wants a spill (host), 2 no parent's children fit
  float fanout;   // children per parent of the last batch, smoothed
  u32 steps;      // batches run since the host last read this record
  u64 states;     // parents taken by those batches
};

struct ExpandArgs {
  const u64* in;        // input states (stride 1 + 6W words); with ctl: the stack, batch at ctl->in_off
  u32 n_in;
  u32* work_next;       // persistent form: next state of the batch to take
  u64* out;             // children
  u32 out_cap;
  u32* out_n;
  u64* retry;           // parents whose children did not fit
  u32* retry_n;
  const PartDesc* parts;
  const u64* rows;
  const u32* maps;
  u32 k;
  int keep;             // write result records
  u32* res;             // result record stream
  u64 res_cap;          // capacity in words
  u64* res_words;       // words used
  u64* res_count;       // results emitted (claims; may exceed limit)
  u64 limit;
  u64* fp;              // [NSTRIPE][2] fingerprint sum / xor
  u64* stat;            // [NSTRIPE][8] states, intersections, leaves, bytes per kernel (3..6)
  u32* flags;
  DedupTable dedup;
  // multi-graph batches: per-graph result caps and digests (null otherwise)
  unsigned long long* gcount;  // results per graph
  u64* gfp;                    // [graph][2] fingerprint sum / xor
  u64 gcap;                    // results kept per graph
  void* scratch;               // wide partitions: one working-set slice per persistent wave
  const LaneCtl* ctl;          // device-planned steps: batch size, position and child room (null: n_in, out_cap)
};

// Seed partitions' vertex sets on the device (seed_kernels.hip).
struct SeedTask {
  u32 graph;   // graph of the batch (index into the bases)
  u32 seed;    // local vertex ID in that graph
  u32 k, lb;
};

struct SeedOut {
  u32 status;  // 0 dropped, 1 built, 2 output arena full (the host builds it), 3 a side above the cap
  u32 nl, nr;  // sides (left, right)
  u32 us;      // side of the seed
  u64 at;      // offset of its lists in the arena: the seed's side, then the other
};

struct SeedArgs {
  const SeedTask* tasks;
  u32 n_tasks;
  const u32* off;        // CSRs of all graphs, concatenated
  const u32* adj;
  const u64* off_base;   // per graph: start of its offsets and lists
  const u64* adj_base;
  const u32* order;      // degeneracy positions of all graphs, at off_base
  const u32* n_vert;     // per graph: vertices
  const u32* n_left;     // per graph: left vertices
  u32 n_max;             // most vertices of any graph (slice stride)
  u32* scratch;          // gridDim.x slices of 8 n_max words (marks, counts, lists, sort)
  u32* arena;
  u64 arena_cap;
  unsigned long long* arena_used;
  SeedOut* out;
  u32 side_cap;          // most vertices per side of a partition (65536)
  unsigned long long* bytes;
};

__device__ __forceinline__ void wave_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ u64 shfl_xor_u64(u64 v, int m) {
  const u32 lo = __shfl_xor((int)(u32)v, m, 64);
  const u32 hi = __shfl_xor((int)(u32)(v >> 32), m, 64);
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 wave_bcast_u64(u64 v, int src) {
  const u32 lo = __shfl((int)(u32)v, src, 64);
  const u32 hi = __shfl((int)(u32)(v >> 32), src, 64);
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 wave_sum_u64(u64 v) {
  for (int m = 32; m > 0; m >>= 1) v += shfl_xor_u64(v, m);
  return v;
}
__device__ __forceinline__ u64 wave_max_u64(u64 v) {
  for (int m = 32; m > 0; m >>= 1) { const u64 o = shfl_xor_u64(v, m); v = o > v ? o : v; }
  return v;
}
__device__ __forceinline__ u64 wave_min_u64(u64 v) {
  for (int m = 32; m > 0; m >>= 1) { const u64 o = shfl_xor_u64(v, m); v = o < v ? o : v; }
  return v;
}

}  // namespace gmk
