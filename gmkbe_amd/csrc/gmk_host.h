// Host-side structures shared by the engine's translation units.
#pragma once
#include <chrono>
#include <string>
#include <vector>

#include "gmk_device.h"
#include "seeds.h"
#include "../../include/gmk.h"

namespace gmk {

// Partitions of one run, with their root search states.
struct PartitionSet {
  int W = 0;                     // bitset words per set (1, 2, 4, ..., 64)
  std::vector<PartDesc> parts;
  std::vector<u64> rows;         // row pool (bitset adjacency)
  std::vector<u32> maps;         // local -> global ID maps
  std::vector<u64> roots;        // root states, 1 + 6W words each
};

struct RunState {
  DedupTable dedup{};
  double dedup_load = 0.5;  // load factor the dedup table is kept under
  bool replay = false;      // produce every result twice (dedup tests)
  double ms_schedule = 0;   // host time spent ordering items (learned runs)
  double ms_reduce = 0;     // core reduction before the seed partitions: device time,
  u64 bytes_reduce = 0;     //   algorithmic bytes,
  u64 reduce_removed = 0;   //   vertices removed and peeling rounds (or passes)
  u64 reduce_edges = 0;     //   edges removed by the butterfly reduction
  double ms_build = 0;      // host time building seed partitions' vertex sets
  u64 reduce_rounds = 0;
  unsigned long long* gcount = nullptr;  // per-graph caps (multi-graph batches)
  u64* gfp = nullptr;
  u64 gcap = ~0ull;
  bool keep = false;
  u64 limit = ~0ull;
  u64 res_cap = 0;
  double ms_expand = 0;
  u64 launches = 0;
  u64 batches = 0;          // branch-and-bound batches planned on the device
  u64 graphs = 0;           // hipGraphs captured for them
  bool hit_limit = false;
  u64 fp_sum_extra = 0, fp_xor_extra = 0;
  u32 spills = 0;
  double ms_kernel[4] = {0, 0, 0, 0};
  u64 launches_kernel[4] = {0, 0, 0, 0};
  // progress log: results so far and elapsed ms after each launch
  std::vector<u64> prog_count;
  std::vector<double> prog_ms;
  double t0_ms = 0;
  void progress(u64 count);
  // verbose runs: a line on stderr at most every 10 s of wall time
  bool verbose = false;
  double last_print_ms = -1e30;
  void report(const char* phase, u64 count, u64 live);
};

// Grows the run's dedup table (doubling, rehashed on the device) so that
// `bound` results fit under its load factor.  Callers pass the results counted
// so far plus the most the next launch can add.
int dedup_reserve(gmk_ctx* c, RunState& rs, u64 bound);

// The loaded graph's (thr, thr)-core, same vertex IDs (removed vertices keep
// empty lists): host copy for the seed construction, device copy (owned by
// the context) for the row builder.
struct CoreGraph {
  std::vector<u32> off, adj;
  const u32* d_off = nullptr;
  const u32* d_adj = nullptr;
  u32 removed = 0;
};
int core_reduce(gmk_ctx* c, u32 thr, RunState& rs, CoreGraph& out);
// Butterfly reduction of a reduced graph's device CSR (vertex IDs kept): the
// vertices and edges in no k-biplex with both sides >= q leave.
int butterfly_reduce(gmk_ctx* c, const CoreGraph& in, u32 q, u32 k, RunState& rs, CoreGraph& out);

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

// State stacks per bitset width 2^l words, l = 0 .. NLANE - 1 (1 .. 1024
// words: partitions of up to 65536 vertices per side).  Widths up to
// LANE_LDS_MAX keep a state's working set in LDS, wider ones in HBM.
constexpr int NLANE = 11;
constexpr int LANE_TUNED = 7;    // widths 1 .. 64: launch forms and batch sizes from the tuning table
constexpr int LANE_LDS_MAX = 7;  // width 128: one workgroup per state with a 100 KB LDS working set

// Device counters of one run, read back after every batch.
struct RunCounters {
  // the run flags, the record-space counter and the result counter sit on
  // cache lines of their own: every emitting workgroup takes record space and
  // result slots with one atomic each, and every wave reads the flags
  alignas(128) u32 flags;
  u32 pad;
  alignas(128) u64 res_words;
  alignas(128) u64 res_count;
  alignas(128) u32 lane[NLANE + 1][4]; // per bitset width 2^l: children, requeued parents, persistent work counter, unused
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
  u32 n_cu = 256;                   // compute units of the device
  // launch table measured on MI355X by gmkbe_amd/autotune.py (gmkbe_amd/tune/mi355x.json)
  gmk_tuning tune{GMK_TUNING_VERSION, {0, 0, 0, 0, 0, 0, 0},
                  {16384, 16384, 16384, 16384, 16384, 16384, 16384}, 500, {0, 0, 0, 0}};
  hipStream_t stream = nullptr;
  std::string err;
  // graph (host copy kept for partition construction)
  u32 n = 0, n_left = 0;
  std::vector<u32> off, adj;
  u32* d_off = nullptr;
  u32* d_adj = nullptr;
  u32* d_deg = nullptr;  // vertex degrees, off[v + 1] - off[v] (one load per lookup)
  // per-run buffers, grown on demand and reused
  DeviceVec lane_stack[NLANE], lane_next[NLANE], lane_retry[NLANE];  // state stacks per bitset width
  DeviceVec lane_spill[NLANE];  // shallow states moved out of a nearly full stack
  DeviceVec wide_scratch;       // working sets of the persistent waves of widths above 128 words
  DeviceVec lane_ctl;           // LaneCtl per width (device-planned steps)
  gmk::LaneCtl* h_ctl = nullptr;  // pinned host copy
  DeviceVec rows, maps, parts, res, counters;
  DeviceVec tuples, arena, fams, prefix, tcnt, tdefer, ridx, tcount;  // thin-result kernels
  DeviceVec nh_off, nh, nh_verts;  // neighbour hashes of the vertices of degree >= nh_min
  u32 nh_min = 0xFFFFFFFFu;        // as built for the current graph (none: all ones)
  u64 nh_slots = 0;
  DeviceVec dedup;                              // result fingerprint table
  DeviceVec gcnt;                               // per-graph counters of a batch
  std::vector<u32> order;                        // degeneracy order (seeded runs)
  bool order_valid = false;
  RunCounters* h_cnt = nullptr;  // pinned
  u64 res_words = 0;             // record stream length of the last run
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  DeviceVec bcsr_off, bcsr_adj, bcsr_base, rowpids;  // CSRs of a batch's graphs; partitions whose rows are built
  std::vector<u64> prog_count;   // progress log of the last run
  std::vector<double> prog_ms;
  bool has_model = false;        // cost model of GMK_MODE_LEARNED batches (per graph)
  gmk_mlp model{};
  bool has_item_model = false;   // cost model of GMK_MODE_LEARNED single-graph runs (per work item)
  gmk_mlp item_model{};
  DeviceVec sched_p, sched_hist, sched_sel;  // learned schedule: model parameters, score keys, selection
  DeviceVec core_deg, core_dead, core_front, core_off, core_adj;  // core reduction of seeded runs
  DeviceVec bfly_cnt, bfly_mark, bfly_deg, bfly_off, bfly_adj;    // butterfly reduction of seeded runs
  DeviceVec seed_order, seed_meta, seed_scratch, seed_arena, seed_tasks, seed_out;  // seed vertex sets on the device
  float* scratch = nullptr;      // device address of the library's static array on this device (gmk_item_predict_static)
  float* h_stage = nullptr;      // pinned host staging of that array's contents
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
namespace gmk {
// Flat parameter vector of the cost model: w1 [in][hid], b1, w2 [hid][hid],
// b2, w3, b3 (the order of gmk_mlp without mu and sd).
constexpr int MLP_NPARAM = GMK_MLP_IN * GMK_MLP_HID + GMK_MLP_HID + GMK_MLP_HID * GMK_MLP_HID + GMK_MLP_HID +
                           GMK_MLP_HID + 1;
void mlp_pack(const gmk_mlp& m, float* P);
// The device address of the static scratch array of gmk_item_predict_static on
// the current device, and the array's size in floats.
int static_scratch(float** ptr, size_t* floats);
// Learned schedule of single-graph runs: about 1.5 `want` one-vertex work
// items of side A ([a_lo, a_hi)) with the highest predicted results per unit
// of device time under the item cost model, sorted best first (ties: lower ID
// first), as (score key << 32 | ~(item - a_lo)) entries (a sampled threshold
// decides which are gathered; the device sorts them).  With pre_max > 0, when
// the front is sure to hold at least pre_max items, the first pre_max of them
// are also written to the device's tuple buffer as one-vertex tuples (pre =
// pre_max) and the front's copy to the host is left in flight: the caller
// takes it with learned_front_finish after its next synchronisation of the
// stream.  Otherwise (pre = 0) the front is in `front` on return.  ms: device
// time of the scoring and selection kernels (known once the front is).
int learned_front(gmk_ctx* c, u32 a_lo, u32 a_hi, u32 k, u32 want, u32 pre_max, std::vector<u64>& front, u32& pre,
                  double& ms);
void learned_front_finish(gmk_ctx* c, std::vector<u64>& front, double& ms);
// Hands out the front's items in order.
struct LearnedOrder {
  std::vector<u64> cand;
  size_t pos = 0;
  bool done() const { return pos >= cand.size(); }
  u32 next(u32 a_lo);
};
}  // namespace gmk
int gmk_predict_costs(gmk_ctx* c, const gmk_graph_desc* graphs, uint32_t n, double* cost);
namespace gmk {
// Seed partitions of one graph, consumed in degeneracy order.
struct SeedJob {
  HostGraph g{};
  const u32* d_off = nullptr;  // the graph's CSR on the device when it is resident there (rows are built from it)
  const u32* d_adj = nullptr;
  u32 k = 1, lb = 3, graph = 0;
  std::vector<u32> order, seeds;
  size_t next = 0;
};
SeedJob make_seed_job(const HostGraph& g, u32 k, u32 lb, u32 graph);
}  // namespace gmk
namespace gmk {
// A partition built on a host thread: offsets relative to its own vectors.
struct BuiltPart {
  int W = 0;
  PartDesc P{};
  bool device_rows = false;  // rows are built on the device from the graph's CSR
  std::vector<u64> rows;     // otherwise: nl rows of wr words, then nr rows of wl words
  std::vector<u32> maps;     // left IDs, then right IDs
  std::vector<u64> root;     // root state without its partition word (6W words)
  u64 row_words() const { return device_rows ? (u64)P.nl * P.wr + (u64)P.nr * P.wl : rows.size(); }
};

// Branch-and-bound over partitions of mixed widths: one device state stack per
// width (1..64 words), one expansion launch per non-empty stack per step, and
// append-only device tables of partitions that may grow while states run.
class BBEngine {
 public:
  BBEngine(gmk_ctx* c, const gmk_params* prm, RunState& rs, bool persistent);
  ~BBEngine();
  BBEngine(const BBEngine&) = delete;
  BBEngine& operator=(const BBEngine&) = delete;
  int reserve(size_t rows_words, size_t maps_words, size_t nparts);
  int make_room(const BuiltPart& b);  // grows empty tables that cannot hold b
  bool fits(const BuiltPart& b) const;
  void add(BuiltPart& b);
  int add_set(const PartitionSet& ps);
  void reset_tables();  // only with no live state
  u64 live() const;     // states on the stacks plus staged roots
  // CSR of every graph a device-built partition may come from (device
  // pointers; per graph the start of its offsets and lists)
  void set_row_source(const u32* off, const u32* adj, const u64* off_base, const u64* adj_base);
  int step();
  int drain();

 private:
  static constexpr u32 kBatchMax = 1u << 16;  // largest states per launch (retry buffer size)
  u32 batch_max(int l) const;                 // states per launch of width 2^l at most
  struct Launch {                             // expansion launch of one width in a step sequence
    int l = 0;
    ExpandArgs a{};
    u32 grid = 0, waves = 0;
  };
  int launch_sequence(const Launch* L, int nl, u32 active, int steps);
  int steps_ = 8;                             // batches per width between host synchronisations
  bool graphs_ = true;                        // replay the step sequence from a hipGraph
  hipGraphExec_t graph_exec_ = nullptr;
  std::vector<char> graph_key_;               // launch parameters of graph_exec_
  int lane_alloc(int l);
  int upload();
  void clear_stage();
  gmk_ctx* c_;
  const gmk_params* prm_;
  RunState& rs_;
  bool persistent_;
  size_t rows_cap_ = 0, maps_cap_ = 0, parts_cap_ = 0;
  size_t rows_used_ = 0, maps_used_ = 0, parts_used_ = 0;  // including staged entries
  size_t rows_dev_ = 0, maps_dev_ = 0, parts_dev_ = 0;     // uploaded
  std::vector<u64> rows_stage_;
  std::vector<u32> maps_stage_;
  std::vector<PartDesc> parts_stage_;
  std::vector<u64> roots_stage_[NLANE], rev_stage_[NLANE];
  std::vector<u32> rowbuild_;  // staged partitions whose rows the device builds
  u32 rowbuild_words_ = 0;     // widest row among them
  bool rows_built_ = false;    // a row build was launched in this step
  const u32 *src_off_ = nullptr, *src_adj_ = nullptr;
  const u64 *src_off_base_ = nullptr, *src_adj_base_ = nullptr;
  u64 staged_roots_ = 0;
  u64 top_[NLANE] = {};
  u64 cap_[NLANE] = {};
  double fanout_[NLANE] = {2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};  // children per parent in the last batch, smoothed
  u64 spill_n_[NLANE] = {};    // states in the spill buffer per width
  u64 spill_cap_[NLANE] = {};  // its capacity in states
  int spill(int l);
  int unspill(int l);
};
}  // namespace gmk

int run_seed_jobs(gmk_ctx* c, std::vector<gmk::SeedJob>& jobs, const gmk_params* prm, gmk_stats* st,
                  gmk::RunState& rs, double* ms_per_job);

int gmk_run_bb(gmk_ctx* c, const gmk::PartitionSet& ps, const gmk_params* prm, gmk_stats* st, gmk::RunState& rs);
int gmk_run_seeded(gmk_ctx* c, const gmk_params* prm, gmk_stats* st, gmk::RunState& rs);
// Thin results with the tuple on side_a (0 left, 1 right); generated side >= min_b.
int gmk_run_thin(gmk_ctx* c, const gmk_params* prm, gmk_stats* st, gmk::RunState& rs, int side_a, uint32_t min_b);
namespace gmk {
// Neighbour hashes of the current graph's vertices of degree >= NH_MIN_DEG
// (the environment variable GMK_NHASH_MIN overrides it; 0: no hashes), used
// by the thin checks' adjacency tests.  Called when a graph is uploaded.
constexpr u32 NH_MIN_DEG = 128;
int build_neighbour_hashes(gmk_ctx* c);
}  // namespace gmk
