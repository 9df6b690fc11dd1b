// Host driver of the branch-and-bound kernels: one device state stack per
// bitset width, drained in batches, with partitions fed while the stacks run.
//
// Seed partitions (BPPivot constructSets, biplex.cpp:315-393) are built on host
// threads only when the stacks run low, graph by graph in schedule order, with
// a per-graph quota that doubles on every visit.  A graph that reaches its
// result cap, or a run that reaches its limit, stops receiving partitions, so a
// capped search builds and expands only the partitions it needs.
#include "gmk_device.h"
#include "gmk_host.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gmk.h"

using namespace gmk;

namespace gmk {
hipError_t launch_bb_expand(int W, u32 n_states, hipStream_t s, const ExpandArgs& a, u32 persistent_waves);
hipError_t launch_bb_plan(LaneCtl* ctl, u32* lane_cnt, const u32* flags, u32 active, int plan, hipStream_t s);
hipError_t launch_bb_commit(const LaneCtl* ctl, const u32* cnt, u64* stack, const u64* next, const u64* retry,
                            u32 stride, hipStream_t s);
size_t bb_wide_scratch_bytes(int W);
struct RowBuildArgs {
  const PartDesc* parts;
  const u32* pids;
  u32 n;
  const u32* maps;
  const u32* off;
  const u32* adj;
  const u64* off_base;
  const u64* adj_base;
  u64* rows;
  u64* bytes;
  u32 max_words;        // widest row of the partitions built
};
hipError_t launch_build_part_rows(const RowBuildArgs& a, hipStream_t s);
hipError_t launch_seed_sets(const SeedArgs& a, u32 grid, hipStream_t s);

namespace {
int lane_of(int W) {
  int l = 0;
  while ((1 << l) < W) ++l;
  return l;
}
// States per launch of widths without a tuning entry (above 64 words)
constexpr u32 kWideBatch = 1024;
// Persistent waves per CU of widths above 128 words (each holds an HBM working set)
constexpr u32 kWideWavesPerCU = 2;
}  // namespace

u32 BBEngine::batch_max(int l) const {
  return l < LANE_TUNED ? std::min<u32>(kBatchMax, c_->tune.batch_states[l]) : kWideBatch;
}

BBEngine::BBEngine(gmk_ctx* c, const gmk_params* prm, RunState& rs, bool persistent)
    : c_(c), prm_(prm), rs_(rs), persistent_(persistent) {
  // GMK_STEPS_PER_SYNC: batches per width between two host synchronisations
  // (1 to 64); GMK_GRAPHS=0: launch them directly instead of from a hipGraph
  if (const char* e = getenv("GMK_STEPS_PER_SYNC")) steps_ = std::clamp(atoi(e), 1, 64);
  if (const char* e = getenv("GMK_GRAPHS")) graphs_ = atoi(e) != 0;
}

// Tables too small for one partition on their own grow, when no state
// references them.
int BBEngine::make_room(const BuiltPart& b) {
  if (fits(b)) return GMK_OK;
  if (live() != 0 || rows_used_ != 0) { c_->err = "partition tables full"; return GMK_EOVERFLOW; }
  return reserve(std::max(rows_cap_, (size_t)b.row_words()), std::max(maps_cap_, b.maps.size()),
                 std::max<size_t>(parts_cap_, 1));
}

int BBEngine::reserve(size_t rows_words, size_t maps_words, size_t nparts) {
  int r;
  if ((r = ensure(c_, c_->rows, std::max<size_t>(8, rows_words * 8)))) return r;
  if ((r = ensure(c_, c_->maps, std::max<size_t>(4, maps_words * 4)))) return r;
  if ((r = ensure(c_, c_->parts, std::max<size_t>(sizeof(PartDesc), nparts * sizeof(PartDesc))))) return r;
  rows_cap_ = c_->rows.bytes / 8;
  maps_cap_ = c_->maps.bytes / 4;
  parts_cap_ = c_->parts.bytes / sizeof(PartDesc);
  return GMK_OK;
}

bool BBEngine::fits(const BuiltPart& b) const {
  return rows_used_ + b.row_words() <= rows_cap_ && maps_used_ + b.maps.size() <= maps_cap_ &&
         parts_used_ + 1 <= parts_cap_;
}

void BBEngine::set_row_source(const u32* off, const u32* adj, const u64* off_base, const u64* adj_base) {
  src_off_ = off;
  src_adj_ = adj;
  src_off_base_ = off_base;
  src_adj_base_ = adj_base;
}

void BBEngine::add(BuiltPart& b) {
  const u32 pid = (u32)parts_used_;
  PartDesc P = b.P;
  P.rows_l += rows_used_;
  P.rows_r += rows_used_;
  P.map_l += maps_used_;
  P.map_r += maps_used_;
  if (b.device_rows) {
    // the row range is reserved here and filled on the device before any state runs
    rowbuild_.push_back(pid);
    rowbuild_words_ = std::max({rowbuild_words_, P.wl, P.wr});
  } else {
    // the stage covers [rows_dev_, rows_used_): zeros over device-built ranges
    // (the build kernel runs after the copy)
    rows_stage_.resize(rows_used_ - rows_dev_, 0);
    rows_stage_.insert(rows_stage_.end(), b.rows.begin(), b.rows.end());
  }
  maps_stage_.insert(maps_stage_.end(), b.maps.begin(), b.maps.end());
  parts_stage_.push_back(P);
  rows_used_ += b.row_words();
  maps_used_ += b.maps.size();
  ++parts_used_;
  std::vector<u64>& rt = roots_stage_[lane_of(b.W)];
  rt.push_back(pid);
  rt.insert(rt.end(), b.root.begin(), b.root.end());
  ++staged_roots_;
}

int BBEngine::add_set(const PartitionSet& ps) {
  int r;
  if (rows_used_ + ps.rows.size() > rows_cap_ || maps_used_ + ps.maps.size() > maps_cap_ ||
      parts_used_ + ps.parts.size() > parts_cap_) {
    if (live() != 0 || rows_used_ != 0) { c_->err = "partition tables full"; return GMK_EOVERFLOW; }
    if ((r = reserve(ps.rows.size(), ps.maps.size(), ps.parts.size()))) return r;
  }
  const u64 base_pid = parts_used_;
  for (PartDesc P : ps.parts) {
    P.rows_l += rows_used_;
    P.rows_r += rows_used_;
    P.map_l += maps_used_;
    P.map_r += maps_used_;
    parts_stage_.push_back(P);
  }
  rows_stage_.resize(rows_used_ - rows_dev_, 0);
  rows_stage_.insert(rows_stage_.end(), ps.rows.begin(), ps.rows.end());
  maps_stage_.insert(maps_stage_.end(), ps.maps.begin(), ps.maps.end());
  rows_used_ += ps.rows.size();
  maps_used_ += ps.maps.size();
  parts_used_ += ps.parts.size();
  const size_t stride = 1 + 6 * (size_t)ps.W;
  std::vector<u64>& rt = roots_stage_[lane_of(ps.W)];
  for (size_t i = 0; i + stride <= ps.roots.size(); i += stride) {
    rt.push_back(ps.roots[i] + base_pid);
    rt.insert(rt.end(), ps.roots.begin() + i + 1, ps.roots.begin() + i + stride);
    ++staged_roots_;
  }
  return GMK_OK;
}

void BBEngine::reset_tables() {
  rows_used_ = maps_used_ = parts_used_ = 0;
  rows_dev_ = maps_dev_ = parts_dev_ = 0;
}

u64 BBEngine::live() const {
  u64 n = staged_roots_;
  for (int l = 0; l < NLANE; ++l) n += top_[l] + spill_n_[l];
  return n;
}

// Moves the bottom (shallowest) ceil(top/2) states of a nearly full stack to the
// width's spill buffer (device memory, grown by doubling with its contents
// copied) and slides the rest down; the stack then has at least half its
// capacity free.
int BBEngine::spill(int l) {
  const u64 stride = 1 + 6ull * (1u << l), sb = stride * 8;
  const u64 move = (top_[l] + 1) / 2;
  if (spill_n_[l] + move > spill_cap_[l]) {
    const u64 ncap = std::max<u64>({cap_[l], 2 * spill_cap_[l], spill_n_[l] + move});
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, ncap * sb);
    if (e != hipSuccess) {
      c_->err = "spill buffer: hipMalloc(" + std::to_string(ncap * sb) + " B) failed: " + hipGetErrorString(e);
      return GMK_ENOMEM;
    }
    if (spill_n_[l])
      HIPCHK(c_, hipMemcpyAsync(p, c_->lane_spill[l].p, spill_n_[l] * sb, hipMemcpyDeviceToDevice, c_->stream));
    HIPCHK(c_, hipStreamSynchronize(c_->stream));
    if (c_->lane_spill[l].p) (void)hipFree(c_->lane_spill[l].p);
    c_->lane_spill[l].p = p;
    c_->lane_spill[l].bytes = ncap * sb;
    spill_cap_[l] = ncap;
  }
  u64* stack = (u64*)c_->lane_stack[l].p;
  u64* sp = (u64*)c_->lane_spill[l].p;
  HIPCHK(c_, hipMemcpyAsync(sp + spill_n_[l] * stride, stack, move * sb, hipMemcpyDeviceToDevice, c_->stream));
  // [move, top) -> [0, top - move): top - move <= move, so the ranges do not overlap
  if (top_[l] > move)
    HIPCHK(c_, hipMemcpyAsync(stack, stack + move * stride, (top_[l] - move) * sb, hipMemcpyDeviceToDevice,
                              c_->stream));
  spill_n_[l] += move;
  top_[l] -= move;
  ++rs_.spills;
  return GMK_OK;
}

// Brings spilled states back into an empty stack (at most half its capacity).
int BBEngine::unspill(int l) {
  const u64 stride = 1 + 6ull * (1u << l), sb = stride * 8;
  const u64 m = std::min<u64>(spill_n_[l], cap_[l] / 2);
  HIPCHK(c_, hipMemcpyAsync(c_->lane_stack[l].p, (u64*)c_->lane_spill[l].p + (spill_n_[l] - m) * stride, m * sb,
                            hipMemcpyDeviceToDevice, c_->stream));
  spill_n_[l] -= m;
  top_[l] = m;
  return GMK_OK;
}

int BBEngine::lane_alloc(int l) {
  if (cap_[l]) return GMK_OK;
  const int W = 1 << l;
  const u64 state_bytes = (1 + 6ull * W) * 8;
  // 4 GiB of states for widths up to 4 words, 8 GiB up to 16, 16 GiB above
  // (HBM3E holds 288 GB: wide searches branch on up to 64 W candidates); the
  // children buffer is half of it
  const u64 bytes = W <= 4 ? (4ull << 30) : W <= 16 ? (8ull << 30) : (16ull << 30);
  // at least 2^17 states, and eight times the most children one state can
  // have (64 W + 1), so a spill always leaves room for the top state's children
  u64 cap = std::max<u64>({1ull << 17, bytes / state_bytes, 8 * (64ull * W + 1)});
  // GMK_STACK_STATES caps every stack (at least 4096 states, and four times
  // the most children one state can have): small stacks exercise the spill
  // path on small graphs
  if (const char* e = getenv("GMK_STACK_STATES")) {
    const u64 v = strtoull(e, nullptr, 10);
    if (v >= 4096) cap = std::min(cap, std::max<u64>(v, 4 * (2ull * 64 * W + 2)));
  }
  int r;
  if ((r = ensure(c_, c_->lane_stack[l], cap * state_bytes))) return r;
  if ((r = ensure(c_, c_->lane_next[l], cap / 2 * state_bytes))) return r;
  if ((r = ensure(c_, c_->lane_retry[l], (u64)batch_max(l) * state_bytes))) return r;
  cap_[l] = cap;
  return GMK_OK;
}

// Uploads staged table entries and pushes staged roots on their stacks (the
// first-fed root on top, expanded first).
int BBEngine::upload() {
  hipStream_t s = c_->stream;
  // Host-built rows are staged contiguously from rows_dev_; device-built rows
  // only reserve their range (rows_stage_ then stays empty, or holds zeros in
  // those ranges when host rows come after them).
  if (!rows_stage_.empty())
    HIPCHK(c_, hipMemcpyAsync((u64*)c_->rows.p + rows_dev_, rows_stage_.data(), rows_stage_.size() * 8,
                              hipMemcpyHostToDevice, s));
  if (!maps_stage_.empty())
    HIPCHK(c_, hipMemcpyAsync((u32*)c_->maps.p + maps_dev_, maps_stage_.data(), maps_stage_.size() * 4,
                              hipMemcpyHostToDevice, s));
  if (!parts_stage_.empty())
    HIPCHK(c_, hipMemcpyAsync((PartDesc*)c_->parts.p + parts_dev_, parts_stage_.data(),
                              parts_stage_.size() * sizeof(PartDesc), hipMemcpyHostToDevice, s));
  rows_dev_ = rows_used_;
  maps_dev_ += maps_stage_.size();
  parts_dev_ += parts_stage_.size();
  if (!rowbuild_.empty()) {
    if (!src_off_) { c_->err = "no graph for device-built partition rows"; return GMK_EINVAL; }
    int r;
    if ((r = ensure(c_, c_->rowpids, rowbuild_.size() * 4))) return r;
    HIPCHK(c_, hipMemcpyAsync(c_->rowpids.p, rowbuild_.data(), rowbuild_.size() * 4, hipMemcpyHostToDevice, s));
    RunCounters* dc = (RunCounters*)c_->counters.p;
    RowBuildArgs ra{(const PartDesc*)c_->parts.p, (const u32*)c_->rowpids.p, (u32)rowbuild_.size(),
                    (const u32*)c_->maps.p, src_off_, src_adj_, src_off_base_, src_adj_base_, (u64*)c_->rows.p,
                    &dc->stat[3 + 3], rowbuild_words_};
    HIPCHK(c_, hipEventRecord(c_->ev2, s));
    HIPCHK(c_, launch_build_part_rows(ra, s));
    HIPCHK(c_, hipEventRecord(c_->ev3, s));
    rows_built_ = true;
  }
  for (int l = 0; l < NLANE; ++l) {
    std::vector<u64>& rt = roots_stage_[l];
    if (rt.empty()) continue;
    int r;
    if ((r = lane_alloc(l))) return r;
    const size_t stride = 1 + 6 * ((size_t)1 << l);
    const size_t n = rt.size() / stride;
    if (top_[l] + n > cap_[l]) { c_->err = "state stack full"; return GMK_EOVERFLOW; }
    // reverse, so that the first staged root ends on top
    std::vector<u64>& rev = rev_stage_[l];
    rev.resize(rt.size());
    for (size_t i = 0; i < n; ++i)
      std::memcpy(rev.data() + (n - 1 - i) * stride, rt.data() + i * stride, stride * 8);
    HIPCHK(c_, hipMemcpyAsync((u64*)c_->lane_stack[l].p + top_[l] * stride, rev.data(), rev.size() * 8,
                              hipMemcpyHostToDevice, s));
    top_[l] += n;
  }
  staged_roots_ = 0;
  return GMK_OK;
}

// Host staging buffers are released after the stream has synchronised.
void BBEngine::clear_stage() {
  rowbuild_.clear();
  rowbuild_words_ = 0;
  rows_stage_.clear();
  maps_stage_.clear();
  parts_stage_.clear();
  for (int l = 0; l < NLANE; ++l) { roots_stage_[l].clear(); rev_stage_[l].clear(); }
}

// One synchronisation of the search: the host first spills nearly full
// stacks and brings spilled states back into empty ones, then the device runs
// steps_ batches per non-empty width, each planned on the device from the
// last one's outcome (bb_plan), expanded (bb_expand) and copied back onto the
// stack (bb_commit), with no host round trip in between.  The sequence is
// captured once into a hipGraph and replayed while its launch parameters stay
// the same (GMK_GRAPHS=0 launches the kernels directly).  Counters are
// cleared by bb_plan, so the graph holds no memset node.
int BBEngine::step() {
  int r;
  if ((r = upload())) return r;
  hipStream_t s = c_->stream;
  RunCounters* dc = (RunCounters*)c_->counters.p;
  u32 active = 0;
  for (int l = 0; l < NLANE; ++l) {
    if (!top_[l] && spill_n_[l] && (r = unspill(l))) return r;
    // Spill when less than a quarter of the stack (and less than twice the most
    // children one state can have) is free: the stack is then at least half
    // free, which the stack's minimum size keeps above the most children of
    // one state, so every batch expands at least its top state.
    if (top_[l] > 1 && cap_[l] - top_[l] < std::max<u64>(cap_[l] / 4, 2 * (2ull * 64 * (1u << l) + 2)) &&
        (r = spill(l)))
      return r;
    if (top_[l]) active |= 1u << l;
  }
  if (!active) return GMK_OK;
  const int S = steps_;
  // each state emits at most one result
  if ((r = dedup_reserve(c_, rs_, c_->h_cnt->res_count + (u64)S * NLANE * kBatchMax))) return r;
  if ((r = ensure(c_, c_->lane_ctl, sizeof(LaneCtl) * NLANE))) return r;
  LaneCtl* dctl = (LaneCtl*)c_->lane_ctl.p;
  LaneCtl* hctl = c_->h_ctl;
  for (int l = 0; l < NLANE; ++l) {
    hctl[l] = LaneCtl{};
    hctl[l].top = top_[l];
    hctl[l].cap = cap_[l];
    hctl[l].bmax = batch_max(l);
    hctl[l].fanout = (float)fanout_[l];
    // The one-workgroup-per-state form launches a workgroup per state of the
    // largest batch the sequence may plan.  For a stack still small at this
    // synchronisation that largest batch is capped at a power of two above
    // 4 times its size, so a small search does not launch the batch cap's
    // worth of workgroups that exit at once.
    if (l < LANE_TUNED && !(persistent_ && c_->tune.waves_per_cu[l])) {
      u64 want = 1024;
      while (want < 4 * top_[l] && want < hctl[l].bmax) want <<= 1;
      hctl[l].bmax = (u32)std::min<u64>(hctl[l].bmax, want);
    }
  }
  HIPCHK(c_, hipMemcpyAsync(dctl, hctl, sizeof(LaneCtl) * NLANE, hipMemcpyHostToDevice, s));
  ExpandArgs a{};
  a.parts = (const PartDesc*)c_->parts.p;
  a.rows = (const u64*)c_->rows.p;
  a.maps = (const u32*)c_->maps.p;
  a.k = prm_->k;
  a.keep = rs_.keep;
  a.res = (u32*)c_->res.p;
  a.res_cap = rs_.res_cap;
  a.res_words = &dc->res_words;
  a.res_count = &dc->res_count;
  a.limit = rs_.limit;
  a.fp = dc->fp;
  a.stat = dc->stat;
  a.flags = &dc->flags;
  a.dedup = rs_.dedup;
  a.gcount = rs_.gcount;
  a.gfp = rs_.gfp;
  a.gcap = rs_.gcap;
  Launch L[NLANE];
  int nl = 0;
  for (int l = 0; l < NLANE; ++l) {
    if (!((active >> l) & 1u)) continue;
    Launch& q = L[nl++];
    q.l = l;
    q.a = a;
    q.a.in = (const u64*)c_->lane_stack[l].p;
    q.a.out = (u64*)c_->lane_next[l].p;
    q.a.retry = (u64*)c_->lane_retry[l].p;
    q.a.out_n = &dc->lane[l][0];
    q.a.retry_n = &dc->lane[l][1];
    q.a.work_next = &dc->lane[l][2];
    q.a.ctl = dctl + l;
    const int W = 1 << l;
    if (l < LANE_TUNED) {
      q.waves = persistent_ ? c_->n_cu * c_->tune.waves_per_cu[l] : 0u;
    } else if (l > LANE_LDS_MAX) {
      // every wave of a wide width works in its own HBM slice
      q.waves = c_->n_cu * kWideWavesPerCU;
      if ((r = ensure(c_, c_->wide_scratch, (size_t)q.waves * bb_wide_scratch_bytes(W)))) return r;
      q.a.scratch = c_->wide_scratch.p;
    }
    // a workgroup per state of the largest batch, or the persistent grid
    q.grid = q.waves ? q.waves : hctl[l].bmax;
  }
  HIPCHK(c_, hipEventRecord(c_->ev0, s));
  if ((r = launch_sequence(L, nl, active, S))) return r;
  HIPCHK(c_, hipEventRecord(c_->ev1, s));
  HIPCHK(c_, hipMemcpyAsync(c_->h_cnt, dc, offsetof(RunCounters, fp), hipMemcpyDeviceToHost, s));
  HIPCHK(c_, hipMemcpyAsync(hctl, dctl, sizeof(LaneCtl) * NLANE, hipMemcpyDeviceToHost, s));
  HIPCHK(c_, hipStreamSynchronize(s));
  clear_stage();
  float ms = 0;
  if (rows_built_) {
    HIPCHK(c_, hipEventElapsedTime(&ms, c_->ev2, c_->ev3));
    rs_.ms_kernel[3] += ms;
    rs_.launches_kernel[3] += 1;
    rs_.launches += 1;
    rows_built_ = false;
  }
  HIPCHK(c_, hipEventElapsedTime(&ms, c_->ev0, c_->ev1));
  rs_.ms_expand += ms;
  rs_.ms_kernel[0] += ms;
  rs_.launches += (u64)S * nl;
  rs_.launches_kernel[0] += (u64)S * nl;
  const RunCounters& h = *c_->h_cnt;
  rs_.progress(std::min<u64>(h.res_count, rs_.limit));
  if (h.flags & FLAG_RES_OVERFLOW) {
    c_->err = "result buffer overflow (raise result_capacity_words)";
    return GMK_EOVERFLOW;
  }
  for (int i = 0; i < nl; ++i) {
    const int l = L[i].l;
    top_[l] = hctl[l].top;
    fanout_[l] = hctl[l].fanout;
    rs_.batches += hctl[l].steps;
    if (hctl[l].halt == 2) {
      // every parent was requeued: the free stack is smaller than the children
      // of each state that ran
      c_->err = "state stack full (no parent's children fit)";
      return GMK_EOVERFLOW;
    }
  }
  if (h.flags & FLAG_LIMIT) rs_.hit_limit = true;
  rs_.report("branch-and-bound", std::min<u64>(h.res_count, rs_.limit), live());
  return GMK_OK;
}

// The kernels of `steps` batches of the widths in L, from a hipGraph when
// graphs are on (re-captured when any launch parameter changed).
int BBEngine::launch_sequence(const Launch* L, int nl, u32 active, int steps) {
  hipStream_t s = c_->stream;
  RunCounters* dc = (RunCounters*)c_->counters.p;
  LaneCtl* dctl = (LaneCtl*)c_->lane_ctl.p;
  auto enqueue = [&]() -> hipError_t {
    hipError_t e = launch_bb_plan(dctl, &dc->lane[0][0], &dc->flags, active, 1, s);
    for (int i = 0; i < steps && e == hipSuccess; ++i) {
      for (int q = 0; q < nl && e == hipSuccess; ++q)
        e = launch_bb_expand(1 << L[q].l, L[q].grid, s, L[q].a, L[q].waves);
      for (int q = 0; q < nl && e == hipSuccess; ++q) {
        const int l = L[q].l;
        e = launch_bb_commit(dctl + l, &dc->lane[l][0], (u64*)c_->lane_stack[l].p, (const u64*)c_->lane_next[l].p,
                             (const u64*)c_->lane_retry[l].p, (u32)(1 + 6 * (1u << l)), s);
      }
      if (e == hipSuccess) e = launch_bb_plan(dctl, &dc->lane[0][0], &dc->flags, active, i + 1 < steps, s);
    }
    return e;
  };
  if (!graphs_) {
    HIPCHK(c_, enqueue());
    return GMK_OK;
  }
  // the graph's key: every launch parameter
  std::vector<char> key;
  auto put = [&](const void* p, size_t n) { key.insert(key.end(), (const char*)p, (const char*)p + n); };
  put(&active, sizeof(active));
  put(&steps, sizeof(steps));
  put(&dctl, sizeof(dctl));
  put(&dc, sizeof(dc));
  for (int q = 0; q < nl; ++q) {
    put(&L[q].a, sizeof(ExpandArgs));
    put(&L[q].grid, sizeof(u32));
    put(&L[q].waves, sizeof(u32));
    const void* bufs[3] = {c_->lane_stack[L[q].l].p, c_->lane_next[L[q].l].p, c_->lane_retry[L[q].l].p};
    put(bufs, sizeof(bufs));
  }
  if (!graph_exec_ || key != graph_key_) {
    if (graph_exec_) { (void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr; }
    hipGraph_t g = nullptr;
    HIPCHK(c_, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const hipError_t e = enqueue();
    const hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      c_->err = std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2);
      return GMK_EDEVICE;
    }
    const hipError_t e3 = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) {
      graph_exec_ = nullptr;
      c_->err = std::string("graph instantiate: ") + hipGetErrorString(e3);
      return GMK_EDEVICE;
    }
    graph_key_ = std::move(key);
    ++rs_.graphs;
  }
  HIPCHK(c_, hipGraphLaunch(graph_exec_, s));
  return GMK_OK;
}

BBEngine::~BBEngine() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
}

int BBEngine::drain() {
  while (live() > 0 && !rs_.hit_limit) {
    int r = step();
    if (r) return r;
  }
  return GMK_OK;
}

// ---------------------------------------------------------------------------
// Partition construction off the main thread.
// ---------------------------------------------------------------------------

static void build_host_rows(const HostGraph& g, const SeedPartition& sp, std::vector<u32>& loc, BuiltPart& b);

// Maps and root state of a seed partition (and its rows unless the device
// builds them), with offsets relative to the part's own vectors.  Returns
// false when a side exceeds 65536 vertices (1024 words).
static bool make_built(const SeedJob& job, const SeedPartition& sp, std::vector<u32>& loc, BuiltPart& b,
                       bool device_rows) {
  const HostGraph& g = job.g;
  const u32 nl = (u32)sp.left.size(), nr = (u32)sp.right.size();
  const u32 wl = (nl + 63) / 64, wr = (nr + 63) / 64;
  const int W = gmk_pick_w(std::max(wl, wr));
  if (!W) return false;
  b.W = W;
  PartDesc& P = b.P;
  P = PartDesc{};
  P.nl = nl; P.nr = nr; P.wl = wl; P.wr = wr;
  P.lb = job.lb;
  P.seed = sp.seed;
  P.k = job.k;
  P.graph = job.graph;
  P.map_l = 0;
  P.map_r = nl;
  P.rows_l = 0;
  P.rows_r = (u64)nl * wr;
  b.maps.resize((size_t)nl + nr);
  std::copy(sp.left.begin(), sp.left.end(), b.maps.begin());
  std::copy(sp.right.begin(), sp.right.end(), b.maps.begin() + nl);
  b.device_rows = device_rows;
  b.rows.clear();
  if (!device_rows) build_host_rows(g, sp, loc, b);
  // root: S = {seed}, C / X from the flags (0 = C, 1 = X, 2 = S)
  b.root.assign(6 * (size_t)W, 0);
  for (int side = 0; side < 2; ++side) {
    const std::vector<uint8_t>& fl = side ? sp.rflag : sp.lflag;
    for (u32 i = 0; i < fl.size(); ++i) {
      const int which = fl[i] == 2 ? 0 : (fl[i] == 0 ? 1 : 2);
      b.root[(side * 3 + which) * W + i / 64] |= 1ull << (i % 64);
    }
  }
  return true;
}

// Host copy of the rows (the whole-graph partition and tests use it).
static void build_host_rows(const HostGraph& g, const SeedPartition& sp, std::vector<u32>& loc, BuiltPart& b) {
  const PartDesc& P = b.P;
  const u32 nl = P.nl, nr = P.nr, wl = P.wl, wr = P.wr;
  for (u32 i = 0; i < nl; ++i) loc[sp.left[i]] = i;
  for (u32 i = 0; i < nr; ++i) loc[sp.right[i]] = i;
  b.rows.assign((size_t)nl * wr + (size_t)nr * wl, 0);
  for (u32 i = 0; i < nl; ++i) {
    const u32 u = sp.left[i];
    u64* row = b.rows.data() + (size_t)i * wr;
    for (u32 e = g.off[u]; e < g.off[u + 1]; ++e) {
      const u32 q = loc[g.adj[e]];
      if (q != NONE && g.adj[e] >= g.n_left) row[q / 64] |= 1ull << (q % 64);
    }
  }
  for (u32 i = 0; i < nr; ++i) {
    const u32 v = sp.right[i];
    u64* row = b.rows.data() + P.rows_r + (size_t)i * wl;
    for (u32 e = g.off[v]; e < g.off[v + 1]; ++e) {
      const u32 q = loc[g.adj[e]];
      if (q != NONE && g.adj[e] < g.n_left) row[q / 64] |= 1ull << (q % 64);
    }
  }
  for (u32 v : sp.left) loc[v] = NONE;
  for (u32 v : sp.right) loc[v] = NONE;
}

namespace {
struct BuildTask {
  u32 job;
  u32 seed;
};
struct BuildOut {
  bool ok = false;       // a partition was built
  bool too_wide = false; // a side exceeds 65536 vertices
  BuiltPart part;
};
struct Worker {
  std::vector<int32_t> cnt;
  std::vector<uint8_t> mark;
  std::vector<u32> loc;
  SeedPartition sp;
};

// Builds the partitions of tasks[i] into outs[i] on up to nthreads threads.
void build_tasks(std::vector<SeedJob>& jobs, const std::vector<BuildTask>& tasks, std::vector<BuildOut>& outs,
                 std::vector<Worker>& workers, bool device_rows) {
  outs.clear();
  outs.resize(tasks.size());
  std::atomic<size_t> next{0};
  auto run = [&](Worker& w) {
    for (;;) {
      const size_t i = next.fetch_add(1);
      if (i >= tasks.size()) break;
      SeedJob& job = jobs[tasks[i].job];
      if (!build_seed_partition(job.g, job.order, tasks[i].seed, job.k, job.lb, w.sp, w.cnt, w.mark)) continue;
      if (make_built(job, w.sp, w.loc, outs[i].part, device_rows)) outs[i].ok = true;
      else outs[i].too_wide = true;
    }
  };
  const size_t nt = std::min(workers.size(), std::max<size_t>(1, tasks.size() / 16));
  if (nt <= 1) {
    run(workers[0]);
    return;
  }
  std::vector<std::thread> th;
  for (size_t t = 1; t < nt; ++t) th.emplace_back(run, std::ref(workers[t]));
  run(workers[0]);
  for (auto& t : th) t.join();
}
}  // namespace

// Seed partitions' vertex sets built on the device (seed_kernels.hip), one
// workgroup per seed; the host then places each partition in the tables as
// before.  A seed whose lists do not fit the output arena is built on the host.
class DeviceSeeds {
 public:
  static constexpr u64 kArenaWords = 1ull << 24;
  static constexpr u32 kSideCap = 65536;
  DeviceSeeds(gmk_ctx* c) : c_(c) {}
  // CSR source (as the row builder's) and the jobs' degeneracy orders, laid out
  // like the offsets (graph g's order at its offsets' base)
  int setup(const std::vector<SeedJob>& jobs, const u32* off, const u32* adj, const u64* off_base,
            const u64* adj_base, const std::vector<u64>& base_h, size_t order_words, u32 ngraphs) {
    off_ = off; adj_ = adj; off_base_ = off_base; adj_base_ = adj_base;
    int r;
    n_max_ = 1;
    std::vector<u32> nv(ngraphs, 0), nl(ngraphs, 0), ord(std::max<size_t>(order_words, 1), 0);
    for (const SeedJob& j : jobs) {
      n_max_ = std::max(n_max_, j.g.n);
      nv[j.graph] = j.g.n;
      nl[j.graph] = j.g.n_left;
      std::copy(j.order.begin(), j.order.end(), ord.begin() + base_h[j.graph]);
    }
    if (n_max_ >= (1u << 30)) return GMK_EUNSUPPORTED;  // the flag bits of the output
    if ((r = ensure(c_, c_->seed_order, ord.size() * 4))) return r;
    if ((r = ensure(c_, c_->seed_meta, (size_t)2 * ngraphs * 4 + 64))) return r;
    grid_ = (u32)std::max<u64>(1, std::min<u64>(2ull * c_->n_cu, (1ull << 30) / (32ull * n_max_)));
    if ((r = ensure(c_, c_->seed_scratch, (size_t)grid_ * 8 * n_max_ * 4))) return r;
    if ((r = ensure(c_, c_->seed_arena, kArenaWords * 4))) return r;
    hipStream_t s = c_->stream;
    HIPCHK(c_, hipMemcpyAsync(c_->seed_order.p, ord.data(), ord.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(c_, hipMemcpyAsync(c_->seed_meta.p, nv.data(), (size_t)ngraphs * 4, hipMemcpyHostToDevice, s));
    HIPCHK(c_, hipMemcpyAsync((u32*)c_->seed_meta.p + ngraphs, nl.data(), (size_t)ngraphs * 4, hipMemcpyHostToDevice, s));
    HIPCHK(c_, hipMemsetAsync(c_->seed_scratch.p, 0, (size_t)grid_ * 8 * n_max_ * 4, s));
    ngraphs_ = ngraphs;
    return GMK_OK;
  }

  int build(std::vector<SeedJob>& jobs, const std::vector<BuildTask>& tasks, std::vector<BuildOut>& outs,
            Worker& w, RunState& rs) {
    outs.clear();
    outs.resize(tasks.size());
    if (tasks.empty()) return GMK_OK;
    hipStream_t s = c_->stream;
    int r;
    const size_t n = tasks.size();
    std::vector<SeedTask> tk(n);
    for (size_t i = 0; i < n; ++i) {
      const SeedJob& j = jobs[tasks[i].job];
      tk[i] = SeedTask{j.graph, tasks[i].seed, j.k, j.lb};
    }
    if ((r = ensure(c_, c_->seed_tasks, n * sizeof(SeedTask) + 64))) return r;
    if ((r = ensure(c_, c_->seed_out, n * sizeof(SeedOut) + 64))) return r;
    unsigned long long* used = (unsigned long long*)((char*)c_->seed_tasks.p + n * sizeof(SeedTask));
    used = (unsigned long long*)(((uintptr_t)used + 15) & ~(uintptr_t)15);
    HIPCHK(c_, hipMemcpyAsync(c_->seed_tasks.p, tk.data(), n * sizeof(SeedTask), hipMemcpyHostToDevice, s));
    HIPCHK(c_, hipMemsetAsync(used, 0, 16, s));
    SeedArgs a{(const SeedTask*)c_->seed_tasks.p, (u32)n, off_, adj_, off_base_, adj_base_, (const u32*)c_->seed_order.p,
               (const u32*)c_->seed_meta.p, (const u32*)c_->seed_meta.p + ngraphs_, n_max_,
               (u32*)c_->seed_scratch.p, (u32*)c_->seed_arena.p, kArenaWords, used, (SeedOut*)c_->seed_out.p,
               kSideCap, (unsigned long long*)&((RunCounters*)c_->counters.p)->stat[3 + 3]};
    HIPCHK(c_, hipEventRecord(c_->ev2, s));
    HIPCHK(c_, launch_seed_sets(a, (u32)std::min<size_t>(grid_, n), s));
    HIPCHK(c_, hipEventRecord(c_->ev3, s));
    std::vector<SeedOut> o(n);
    unsigned long long u[2] = {0, 0};
    HIPCHK(c_, hipMemcpyAsync(o.data(), c_->seed_out.p, n * sizeof(SeedOut), hipMemcpyDeviceToHost, s));
    HIPCHK(c_, hipMemcpyAsync(u, used, 16, hipMemcpyDeviceToHost, s));
    HIPCHK(c_, hipStreamSynchronize(s));
    const u64 words = std::min<u64>(u[0], kArenaWords);
    arena_.resize(std::max<u64>(words, 1));
    if (words) {
      HIPCHK(c_, hipMemcpyAsync(arena_.data(), c_->seed_arena.p, words * 4, hipMemcpyDeviceToHost, s));
      HIPCHK(c_, hipStreamSynchronize(s));
    }
    float ms = 0;
    HIPCHK(c_, hipEventElapsedTime(&ms, c_->ev2, c_->ev3));
    rs.ms_kernel[3] += ms;  // partition construction
    rs.launches_kernel[3] += 1;
    rs.launches += 1;
    for (size_t i = 0; i < n; ++i) {
      SeedJob& job = jobs[tasks[i].job];
      BuildOut& bo = outs[i];
      if (o[i].status == 0) continue;
      if (o[i].status == 3) { bo.too_wide = true; continue; }
      if (o[i].status == 2) {
        // output arena full: this seed is built on the host
        if (!build_seed_partition(job.g, job.order, tasks[i].seed, job.k, job.lb, w.sp, w.cnt, w.mark)) continue;
      } else {
        SeedPartition& sp = w.sp;
        sp.seed = tasks[i].seed;
        sp.us = (int)o[i].us;
        const u32 ns = o[i].us ? o[i].nr : o[i].nl, no = o[i].us ? o[i].nl : o[i].nr;
        const u32* same = arena_.data() + o[i].at;
        const u32* other = same + ns;
        std::vector<u32>& vs = sp.us ? sp.right : sp.left;
        std::vector<uint8_t>& fs = sp.us ? sp.rflag : sp.lflag;
        std::vector<u32>& vo = sp.us ? sp.left : sp.right;
        std::vector<uint8_t>& fo = sp.us ? sp.lflag : sp.rflag;
        vs.resize(ns);
        fs.resize(ns);
        for (u32 q = 0; q < ns; ++q) { vs[q] = same[q] & 0x3FFFFFFFu; fs[q] = (uint8_t)(same[q] >> 30); }
        vo.assign(other, other + no);
        fo.assign(no, 0);
      }
      if (make_built(job, w.sp, w.loc, bo.part, true)) bo.ok = true;
      else bo.too_wide = true;
    }
    return GMK_OK;
  }

 private:
  gmk_ctx* c_;
  const u32 *off_ = nullptr, *adj_ = nullptr;
  const u64 *off_base_ = nullptr, *adj_base_ = nullptr;
  u32 n_max_ = 1, grid_ = 1, ngraphs_ = 1;
  std::vector<u32> arena_;
};

}  // namespace gmk

// Runs the seed partitions of the jobs (seeds of each job in degeneracy order).
//   plain:    one partition at a time, each drained before the next is built
//             (the reference's per-partition processing); jobs in order, and
//             ms_per_job (optional) receives each job's device time;
//   batched:  all jobs share the stacks; partitions are built on host threads
//             while the stacks run, jobs in order with a doubling quota.
int run_seed_jobs(gmk_ctx* c, std::vector<SeedJob>& jobs, const gmk_params* prm, gmk_stats* st, RunState& rs,
                  double* ms_per_job) {
  if (jobs.empty()) return GMK_OK;  // (a batch whose graphs all run as global partitions)
  const bool plain = prm->mode == GMK_MODE_PLAIN;
  std::vector<BuildTask> tasks;
  std::vector<BuildOut> outs;
  BBEngine eng(c, prm, rs, !plain);
  int r;
  if ((r = eng.reserve((size_t)1 << 27, (size_t)1 << 25, (size_t)1 << 21))) return r;  // 1 GiB of rows
  u32 nmax = 0, ngraphs = 0;
  for (auto& j : jobs) { nmax = std::max(nmax, j.g.n); ngraphs = std::max(ngraphs, j.graph + 1); }
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  std::vector<Worker> workers(plain ? 1 : std::min(16u, hw));
  for (Worker& w : workers) {
    w.cnt.assign(nmax, 0);
    w.mark.assign(nmax, 0);
    w.loc.assign(nmax, NONE);
  }
  // CSR of every job's graph for the device row builder: the context's own
  // graph, or the batch's graphs concatenated
  std::vector<u64> base(2 * (size_t)ngraphs, 0);
  std::vector<u32> coff, cadj;
  // one resident graph (the context's, or its core): rows come from its device CSR
  bool resident = jobs[0].d_off != nullptr;
  for (auto& j : jobs) resident = resident && j.d_off == jobs[0].d_off && j.graph == 0;
  if (resident) {
    if ((r = ensure(c, c->bcsr_base, 16))) return r;
    HIPCHK(c, hipMemsetAsync(c->bcsr_base.p, 0, 16, c->stream));
    eng.set_row_source(jobs[0].d_off, jobs[0].d_adj, (const u64*)c->bcsr_base.p, (const u64*)c->bcsr_base.p + 1);
  } else {
    std::vector<char> seen(ngraphs, 0);
    for (auto& j : jobs) {
      if (seen[j.graph]) continue;
      seen[j.graph] = 1;
      base[j.graph] = coff.size();
      base[ngraphs + j.graph] = cadj.size();
      coff.insert(coff.end(), j.g.off, j.g.off + j.g.n + 1);
      cadj.insert(cadj.end(), j.g.adj, j.g.adj + j.g.off[j.g.n]);
    }
    if ((r = ensure(c, c->bcsr_off, std::max<size_t>(4, coff.size() * 4)))) return r;
    if ((r = ensure(c, c->bcsr_adj, std::max<size_t>(4, cadj.size() * 4)))) return r;
    if ((r = ensure(c, c->bcsr_base, base.size() * 8))) return r;
    HIPCHK(c, hipMemcpyAsync(c->bcsr_off.p, coff.data(), coff.size() * 4, hipMemcpyHostToDevice, c->stream));
    if (!cadj.empty())
      HIPCHK(c, hipMemcpyAsync(c->bcsr_adj.p, cadj.data(), cadj.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->bcsr_base.p, base.data(), base.size() * 8, hipMemcpyHostToDevice, c->stream));
    eng.set_row_source((const u32*)c->bcsr_off.p, (const u32*)c->bcsr_adj.p, (const u64*)c->bcsr_base.p,
                       (const u64*)c->bcsr_base.p + ngraphs);
  }
  // vertex sets on the device (GMK_DEVICE_SEEDS=0: on host threads)
  DeviceSeeds dev(c);
  bool dev_on = true;
  if (const char* e = getenv("GMK_DEVICE_SEEDS")) dev_on = atoi(e) != 0;
  if (dev_on) {
    const int rr = resident ? dev.setup(jobs, jobs[0].d_off, jobs[0].d_adj, (const u64*)c->bcsr_base.p,
                                        (const u64*)c->bcsr_base.p + 1, std::vector<u64>(2, 0), jobs[0].g.n + 1, 1)
                            : dev.setup(jobs, (const u32*)c->bcsr_off.p, (const u32*)c->bcsr_adj.p,
                                        (const u64*)c->bcsr_base.p, (const u64*)c->bcsr_base.p + ngraphs, base,
                                        coff.size(), ngraphs);
    if (rr == GMK_EUNSUPPORTED) dev_on = false;
    else if (rr) return rr;
  }
  auto build = [&]() -> int {
    if (dev_on) return dev.build(jobs, tasks, outs, workers[0], rs);
    build_tasks(jobs, tasks, outs, workers, true);
    return GMK_OK;
  };
  std::vector<u64> gcnt_h(ngraphs, 0);
  auto capped = [&](const SeedJob& j) { return rs.gcount && gcnt_h[j.graph] >= rs.gcap; };
  auto done = [&](const SeedJob& j) { return j.next >= j.seeds.size() || capped(j); };
  auto too_wide = [&](u32 seed) {
    c->err = "seed partition of vertex " + std::to_string(seed) + " has a side above 65536 vertices";
    return GMK_EUNSUPPORTED;
  };

  if (plain) {
    for (size_t ji = 0; ji < jobs.size() && !rs.hit_limit; ++ji) {
      SeedJob& j = jobs[ji];
      const double t0 = rs.ms_kernel[0];
      while (!done(j) && !rs.hit_limit) {
        tasks.assign(1, BuildTask{(u32)ji, j.seeds[j.next++]});
        const auto tb = std::chrono::steady_clock::now();
        if ((r = build())) return r;
        rs.ms_build += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tb).count();
        if (outs[0].too_wide) return too_wide(tasks[0].seed);
        if (!outs[0].ok) continue;
        ++st->partitions;
        st->widest_side = std::max<u64>({st->widest_side, outs[0].part.P.nl, outs[0].part.P.nr});
        if ((r = eng.make_room(outs[0].part))) return r;
        eng.add(outs[0].part);
        if ((r = eng.drain())) return r;
        eng.reset_tables();
        if (rs.gcount) {
          HIPCHK(c, hipMemcpyAsync(&gcnt_h[j.graph], rs.gcount + j.graph, 8, hipMemcpyDeviceToHost, c->stream));
          HIPCHK(c, hipStreamSynchronize(c->stream));
        }
      }
      if (ms_per_job) ms_per_job[ji] = rs.ms_kernel[0] - t0;
    }
    return GMK_OK;
  }

  const u64 low_water = 8192;  // feed when fewer states than this are live
  const size_t want = 2048;    // seeds examined per feed
  std::vector<u32> quota(jobs.size(), 32);
  std::deque<BuiltPart> carry;  // built, waiting for table room
  size_t cur = 0;               // jobs before cur are done
  for (;;) {
    if (rs.hit_limit) break;
    while (!carry.empty() && eng.fits(carry.front())) {
      eng.add(carry.front());
      carry.pop_front();
    }
    if (carry.empty() && eng.live() < low_water) {
      if (rs.gcount) {
        HIPCHK(c, hipMemcpyAsync(gcnt_h.data(), rs.gcount, (size_t)ngraphs * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
      }
      while (cur < jobs.size() && done(jobs[cur])) ++cur;
      tasks.clear();
      for (size_t ji = cur; ji < jobs.size() && tasks.size() < want; ++ji) {
        SeedJob& j = jobs[ji];
        if (done(j)) continue;
        const size_t take = std::min<size_t>({(size_t)quota[ji], j.seeds.size() - j.next, want - tasks.size()});
        for (size_t q = 0; q < take; ++q) tasks.push_back(BuildTask{(u32)ji, j.seeds[j.next++]});
        quota[ji] = std::min<u32>(quota[ji] * 2, 1u << 16);
      }
      if (!tasks.empty()) {
        const auto tb = std::chrono::steady_clock::now();
        if ((r = build())) return r;
        rs.ms_build += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tb).count();
      }
      for (size_t i = 0; i < tasks.size(); ++i) {
        BuildOut& o = outs[i];
        if (o.too_wide) return too_wide(tasks[i].seed);
        if (!o.ok) continue;
        ++st->partitions;
        st->widest_side = std::max<u64>({st->widest_side, o.part.P.nl, o.part.P.nr});
        if (carry.empty() && eng.fits(o.part)) eng.add(o.part);
        else carry.push_back(std::move(o.part));
      }
    }
    if (eng.live() == 0) {
      eng.reset_tables();  // nothing references the tables any more
      if (!carry.empty()) {
        if ((r = eng.make_room(carry.front()))) return r;
        continue;
      }
      bool more = false;
      for (size_t ji = cur; ji < jobs.size() && !more; ++ji) more = !done(jobs[ji]);
      if (!more) break;
      continue;
    }
    if ((r = eng.step())) return r;
  }
  return GMK_OK;
}

// Whole partition set, all roots at once.
int gmk_run_bb(gmk_ctx* c, const PartitionSet& ps, const gmk_params* prm, gmk_stats* st, RunState& rs) {
  for (const PartDesc& P : ps.parts) st->widest_side = std::max<u64>({st->widest_side, P.nl, P.nr});
  BBEngine eng(c, prm, rs, prm->mode != GMK_MODE_PLAIN);
  int r;
  if ((r = eng.reserve(ps.rows.size(), ps.maps.size(), ps.parts.size()))) return r;
  if ((r = eng.add_set(ps))) return r;
  return eng.drain();
}
