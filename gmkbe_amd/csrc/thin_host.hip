// Host orchestration of the thin-result kernels: tuples are generated seed by
// seed (ascending IDs on the tuple side), families built on the device in
// chunks, then every concrete combination is checked on the device, rank range
// by rank range (each range: rank index, first pass, deferred pass), until the
// result limit is reached.  Also builds the graph's neighbour hashes at upload.
#include <algorithm>
#include <chrono>
#include <cstddef>
#include <functional>
#include <vector>

#include "gmk_host.h"
#include "thin_args.h"

using namespace gmk;
using thin::Family;

namespace {

struct ThinCounters {
  u32 fam_n, tflags;
  u64 arena_used;
  u32 work, defer_n, dwork, pad;  // check kernels: chunk counter, deferred ranks, next deferred rank
  u64 total;                      // ranks of the first slice of the launch's families
};

constexpr u32 TCH = 1u << 16;          // tuples per family-construction launch
constexpr u64 ARENA = 1ull << 27;      // arena words (512 MiB)
constexpr u32 FCAP = 1u << 18;         // families per launch
constexpr u64 RCH = 1ull << 22;        // ranks per check launch
constexpr u64 SHRINK_MIN = 1ull << 16;  // remaining results below which a launch takes the slot check
constexpr u32 DEFER_BUDGET = 64;       // candidates one lane tests before a rank goes to a whole wave
constexpr u32 LEARNED_FRONT = 4096;    // about 1.5 times this many one-vertex items are ordered by the learned schedule (the rest follow in ID order)
constexpr u64 LEARNED_SLICE = 1ull << 24; // ranks of each family per slice in learned runs

struct ThinRun {
  gmk_ctx* c;
  const gmk_params* prm;
  RunState* rs;
  ThinArgs a{};
  std::vector<u64> totals, prefix;
  bool stop = false;
  // probe runs (gmk_probe_items): ranks checked per family, and per item
  // (vertex - probe_first) the results and lane ticks
  u64 probe_cap = ~0ull;
  u64 slice = 0;  // learned runs: ranks of each family per slice (0: whole families)
  u32 probe_first = 0;
  std::vector<u64>* probe_res = nullptr;
  std::vector<u64>* probe_ticks = nullptr;
  bool guards = true;  // guard tables for the checks (GMK_GUARDS=0: plain checks only)
  u64 count0 = 0;          // result count when the run's thin part started
  u64 ranks_settled = 0;   // combinations launched since then whose results the host has read
};

// Checks ranks [0, total) of the current family table (its prefix sums are
// on the device), RCH ranks per launch: rank index, first pass, deferred pass.
// Launches go out without waiting for the ones before them while the count
// they may reach stays under the limit (each rank yields at most one result):
// the host reads the counters after a launch that may reach the limit, after
// MAX_FLIGHT launches, and at the end of the range.
constexpr u32 MAX_FLIGHT = 4;
int check_ranks(ThinRun& T, u64 total) {
  gmk_ctx* c = T.c;
  hipStream_t s = c->stream;
  ThinArgs& a = T.a;
  RunCounters* dc = (RunCounters*)c->counters.p;
  const u64 limit = T.rs->limit;
  u64 inflight = 0;  // ranks launched since the counters were last read
  u32 nflight = 0;
  // waits for the launches in flight and reads the counters
  auto settle = [&]() -> int {
    HIPCHK(c, hipEventRecord(c->ev1, s));
    HIPCHK(c, hipMemcpyAsync(c->h_cnt, dc, offsetof(RunCounters, fp), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    T.rs->ms_expand += ms;
    T.rs->ms_kernel[2] += ms;
    T.ranks_settled += inflight;
    inflight = 0;
    nflight = 0;
    T.rs->progress(std::min<u64>(c->h_cnt->res_count, T.rs->limit));
    T.rs->report(a.swap_sides ? "thin (tuples on the right)" : "thin (tuples on the left)",
                 std::min<u64>(c->h_cnt->res_count, T.rs->limit), 0);
    if (c->h_cnt->flags & FLAG_RES_OVERFLOW) { c->err = "result buffer overflow (raise result_capacity_words)"; return GMK_EOVERFLOW; }
    if (c->h_cnt->flags & FLAG_LIMIT) { T.stop = true; T.rs->hit_limit = true; return GMK_OK; }
    if (c->h_cnt->res_count >= limit) {
      // unlimited launches ended exactly at the limit: raise the flag the
      // kernels of later phases read
      c->h_cnt->flags |= FLAG_LIMIT;
      HIPCHK(c, hipMemcpyAsync(&dc->flags, &c->h_cnt->flags, sizeof(u32), hipMemcpyHostToDevice, s));
      HIPCHK(c, hipStreamSynchronize(s));
      T.stop = true;
      T.rs->hit_limit = true;
    }
    return GMK_OK;
  };
  // GMK_DEDUP_REPLAY: every rank range a second time (its results repeat)
  const u64 passes = T.rs->replay ? 2 : 1;
  for (u64 pass = 0, r0 = 0; pass < passes; r0 = 0, ++pass)
  for (; r0 < total; r0 = a.r1) {
    // The count when this launch starts is at most cnt + inflight.  A launch
    // that cannot pass the limit from there skips the per-result slot check
    // (unlimited).  Near the limit, while the run's launches have been
    // yielding at least half a result per rank, a launch shrinks to what
    // remains (then covering at least half of it); below SHRINK_MIN remaining
    // results, or at a lower yield, a launch takes the slot check.
    const u64 cnt = c->h_cnt->res_count;
    const u64 bound = cnt + inflight;
    u64 n = std::min(RCH, total - r0);
    if (limit != ~0ull) {
      const u64 rem = limit > bound ? limit - bound : 0;
      const bool high_yield = T.ranks_settled && 2 * (cnt - T.count0) >= T.ranks_settled;
      if (rem < n && rem >= SHRINK_MIN && high_yield) n = rem;
    }
    a.r0 = r0;
    a.r1 = r0 + n;
    a.unlimited = limit == ~0ull || bound + n <= limit;
    int rr;
    if ((rr = dedup_reserve(c, *T.rs, bound + n))) return rr;
    a.dedup = T.rs->dedup;
    if (nflight == 0) HIPCHK(c, hipEventRecord(c->ev0, s));
    // the rank index also clears work, defer_n and dwork (consecutive)
    HIPCHK(c, launch_thin_rank_index((const u64*)c->prefix.p, a.nfam, a.r0, a.r1, (u32*)c->ridx.p,
                                     &((ThinCounters*)c->tcnt.p)->work, s));
    HIPCHK(c, launch_thin_check(a, s, c->n_cu * 8, c->tune.check_form[a.k < 4 ? a.k : 3]));
    HIPCHK(c, launch_thin_check_deferred(a, s, c->n_cu * 2));
    T.rs->launches_kernel[2]++;
    T.rs->launches++;
    inflight += n;
    ++nflight;
    if (!a.unlimited || nflight >= MAX_FLIGHT || a.r1 >= total) {
      if ((rr = settle())) return rr;
      if (T.stop) return GMK_OK;
    }
  }
  return GMK_OK;
}

// One launch of `count` tuples of j vertices: families, guard tables, then the
// checks of their ranks.  With on_device the tuples are already in the
// device's tuple buffer, and `tuples` (their host copy) is filled by
// after_sync, which runs after the launch's first synchronisation.
int run_chunk(ThinRun& T, const u32* tuples, u32 count, u32 j, const std::function<void()>* after_sync = nullptr,
              bool on_device = false) {
  gmk_ctx* c = T.c;
  hipStream_t s = c->stream;
  if (!on_device)
    HIPCHK(c, hipMemcpyAsync(c->tuples.p, tuples, (size_t)count * j * 4, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemsetAsync(c->tcnt.p, 0, sizeof(ThinCounters), s));
  ThinArgs& a = T.a;
  a.ntuples = count;
  a.j = j;
  HIPCHK(c, hipEventRecord(c->ev0, s));
  HIPCHK(c, launch_thin_build(a, s));
  if (T.guards) HIPCHK(c, launch_thin_guard(a, s, c->n_cu * 4));
  HIPCHK(c, hipEventRecord(c->ev1, s));
  // the first rank slice's prefix sums, over the family count on the device,
  // so that one read brings the count and the slice's total
  ThinCounters* dtc = (ThinCounters*)c->tcnt.p;
  HIPCHK(c, launch_thin_scan_dev((const Family*)c->fams.p, &dtc->fam_n, FCAP, (u64*)c->prefix.p, 0,
                                 T.slice ? T.slice : T.probe_cap, &dtc->total, s));
  ThinCounters tc;
  HIPCHK(c, hipMemcpyAsync(&tc, c->tcnt.p, sizeof(tc), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (after_sync) (*after_sync)();
  {
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    T.rs->ms_expand += ms;
    T.rs->ms_kernel[1] += ms;
    T.rs->launches_kernel[1]++;
  }
  T.rs->launches++;
  T.rs->report(a.swap_sides ? "thin families (tuples on the right)" : "thin families (tuples on the left)",
               std::min<u64>(c->h_cnt->res_count, T.rs->limit), 0);
  if (tc.tflags) {
    if (count == 1) { c->err = "one tuple's neighbour lists exceed the arena"; return GMK_EOVERFLOW; }
    const u32 h = count / 2;
    int r = run_chunk(T, tuples, h, j);
    if (r || T.stop) return r;
    return run_chunk(T, tuples + (size_t)h * j, count - h, j);
  }
  const u32 nfam = tc.fam_n;
  if (nfam == 0) return GMK_OK;
  a.nfam = nfam;
  // Rank slices: learned runs take every family's ranks [base, base + slice)
  // in turn, the families in schedule order within each slice; other runs
  // take each family whole (probes: its first probe_cap ranks) in one slice.
  for (u64 base = 0;; base += T.slice) {
    u64 total = tc.total;
    if (base) {
      HIPCHK(c, launch_thin_scan((const Family*)c->fams.p, nfam, (u64*)c->prefix.p, base, T.slice, s));
      HIPCHK(c, hipMemcpyAsync(&total, (u64*)c->prefix.p + nfam, 8, hipMemcpyDeviceToHost, s));
      HIPCHK(c, hipStreamSynchronize(s));
    }
    if (total == 0) break;
    a.rank_base = base;
    const int r = check_ranks(T, total);
    if (r || T.stop) return r;
    if (!T.slice) break;
  }
  if (a.item_log) {
    // probe: this launch's per-tuple log, by item
    std::vector<u64> lg(2ull * count);
    HIPCHK(c, hipMemcpyAsync(lg.data(), a.item_log, lg.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemsetAsync(a.item_log, 0, lg.size() * 8, s));
    HIPCHK(c, hipStreamSynchronize(s));
    for (u32 t = 0; t < count; ++t) {
      const u32 it = tuples[(size_t)t * j] - T.probe_first;
      (*T.probe_res)[it] += lg[2 * t];
      (*T.probe_ticks)[it] += lg[2 * t + 1];
    }
  }
  return GMK_OK;
}

// Kernel arguments and buffers of a thin run with the tuple on side_a.
int thin_setup(gmk_ctx* c, const gmk_params* prm, RunState& rs, int side_a, uint32_t min_b, ThinRun& T) {
  const u32 k = prm->k;
  int r;
  if ((r = ensure(c, c->tuples, (size_t)TCH * thin::MAXJ * 4))) return r;
  if ((r = ensure(c, c->arena, ARENA * 4))) return r;
  if ((r = ensure(c, c->fams, (size_t)FCAP * sizeof(Family)))) return r;
  if ((r = ensure(c, c->prefix, (size_t)(FCAP + 1) * 8))) return r;
  if ((r = ensure(c, c->tcnt, sizeof(ThinCounters)))) return r;
  if ((r = ensure(c, c->tdefer, RCH * 8))) return r;
  if ((r = ensure(c, c->ridx, ((RCH >> RIDX_LOG) + 2) * 4))) return r;
  if ((r = ensure(c, c->tcount, (size_t)TCH * 8 * 4))) return r;
  T.c = c; T.prm = prm; T.rs = &rs;
  T.count0 = c->h_cnt->res_count;
  ThinArgs& a = T.a;
  a.g.off = c->d_off;
  a.g.adj = c->d_adj;
  a.g.nh_off = (const u64*)c->nh_off.p;
  a.g.nh = c->nh_slots ? (const u32*)c->nh.p : nullptr;
  a.g.nh_min = c->nh_min;
  if (side_a == 0) { a.g.a_lo = 0; a.g.a_hi = c->n_left; a.g.b_lo = c->n_left; a.g.b_hi = c->n; }
  else { a.g.a_lo = c->n_left; a.g.a_hi = c->n; a.g.b_lo = 0; a.g.b_hi = c->n_left; }
  a.k = k;
  a.min_b = min_b;
  a.tuples = (const u32*)c->tuples.p;
  a.arena = (u32*)c->arena.p;
  a.arena_cap = ARENA;
  ThinCounters* tc = (ThinCounters*)c->tcnt.p;
  a.arena_used = &tc->arena_used;
  a.fams = (Family*)c->fams.p;
  a.fam_cap = FCAP;
  a.fam_n = &tc->fam_n;
  a.tflags = &tc->tflags;
  a.work = &tc->work;
  a.budget = DEFER_BUDGET;
  // GMK_CHECK_BUDGET sets the candidates one lane tests before deferring (0
  // defers every open rank): small budgets exercise the deferred pass on
  // small graphs
  if (const char* e = getenv("GMK_CHECK_BUDGET")) a.budget = (u32)std::min<unsigned long>(strtoul(e, nullptr, 10), 1u << 20);
  // GMK_WAVE1_TUPLES overrides the launch size below which one-vertex tuples
  // are built a wave each (0: always a lane each), so tests cover both forms
  a.wave1_max = WAVE1_MAX_TUPLES;
  if (const char* e = getenv("GMK_WAVE1_TUPLES")) a.wave1_max = (u32)std::min<unsigned long>(strtoul(e, nullptr, 10), ~0u);
  // GMK_GUARDS=0 turns the families' guard tables off (every rank takes the
  // plain candidate lists), so tests can compare the two checks
  if (const char* e = getenv("GMK_GUARDS")) T.guards = atoi(e) != 0;
  // GMK_GUARD_LISTS=0 keeps the guard tables for whole-family decisions only
  a.guard_lists = 1;
  if (const char* e = getenv("GMK_GUARD_LISTS")) a.guard_lists = atoi(e) != 0;
  // GMK_GUARD_MIN_RANKS overrides the family size below which no table is
  // built (0: every family), so small test graphs exercise the tables
  a.guard_min_ranks = thin::GUARD_MIN_RANKS;
  if (const char* e = getenv("GMK_GUARD_MIN_RANKS")) a.guard_min_ranks = strtoull(e, nullptr, 10);
  a.defer = (u64*)c->tdefer.p;
  a.ridx = (const u32*)c->ridx.p;
  a.tcount = (u32*)c->tcount.p;
  a.defer_n = &tc->defer_n;
  a.dwork = &tc->dwork;
  RunCounters* dc = (RunCounters*)c->counters.p;
  a.flags = &dc->flags;
  a.prefix = (const u64*)c->prefix.p;
  a.keep = rs.keep;
  a.swap_sides = side_a;
  a.res = (u32*)c->res.p;
  a.res_cap = rs.res_cap;
  a.res_words = &dc->res_words;
  a.res_count = &dc->res_count;
  a.limit = rs.limit;
  a.fp = dc->fp;
  a.stat = dc->stat;
  a.dedup = rs.dedup;
  a.item_log = nullptr;
  return GMK_OK;
}

// Streams tuples into launches: `count` tuples of j vertices built by
// `next(idx)` (false at the end), launch sizes ramping from 64 tuples (x4 per
// launch) to TCH so a run that reaches its limit within the first tuples stops
// early; `per_seed` flushes after every first vertex.
struct TupleFeed {
  ThinRun& T;
  u32 j;
  u32& chunk;
  std::vector<u32> buf;
  u32 count = 0;
  u64 est = 0;
  TupleFeed(ThinRun& t, u32 jj, u32& ch) : T(t), j(jj), chunk(ch) { buf.reserve((size_t)TCH * thin::MAXJ); }
  int flush() {
    if (count == 0) return GMK_OK;
    const int rr = run_chunk(T, buf.data(), count, j);
    buf.clear(); count = 0; est = 0;
    chunk = std::min<u32>(chunk * 4, TCH);
    return rr;
  }
  int push(const u32* idx) {
    const u32* off = T.c->off.data();
    u64 need = 0;
    for (u32 i = 0; i < j; ++i) need += 2ull * (off[idx[i] + 1] - off[idx[i]]);
    if (count >= chunk || est + need > ARENA) {
      const int rr = flush();
      if (rr || T.stop) return rr;
    }
    buf.insert(buf.end(), idx, idx + j);
    ++count;
    est += need;
    return GMK_OK;
  }
};

}  // namespace

int gmk_run_thin(gmk_ctx* c, const gmk_params* prm, gmk_stats* st, RunState& rs, int side_a, uint32_t min_b) {
  (void)st;
  const u32 k = prm->k;
  const u32 jmin = std::max<u32>(prm->theta, 1);
  const u32 jmax = std::min<u32>(2 * k, thin::MAXJ);
  if (jmin > jmax) return GMK_OK;
  ThinRun T;
  int r;
  if ((r = thin_setup(c, prm, rs, side_a, min_b, T))) return r;
  const u32 a_lo = T.a.g.a_lo, a_hi = T.a.g.a_hi;
  const bool per_seed = prm->mode == GMK_MODE_PLAIN;
  u32 chunk = 64;
  u32 jfirst = jmin;
  // Learned order (GMK_MODE_LEARNED with an item model): the one-vertex items
  // with the best predicted results per unit of device time come first, then
  // every other one-vertex item in ID order; tuples of two or more vertices
  // follow in blocks as below.
  if (prm->mode == GMK_MODE_LEARNED && c->has_item_model && jmin == 1) {
    T.slice = LEARNED_SLICE;
    // GMK_LEARNED_SLICE overrides the ranks per family per slice (at least 1024)
    if (const char* e = getenv("GMK_LEARNED_SLICE")) T.slice = std::max<u64>(1024, strtoull(e, nullptr, 10));
    LearnedOrder order;
    u32 pre = 0;
    double ms_sched = 0;
    if ((r = learned_front(c, a_lo, a_hi, k, LEARNED_FRONT, chunk, order.cand, pre, ms_sched))) return r;
    std::vector<u32> taken;
    if (pre) {
      // The front's first `chunk` items are the first launch's tuples, on the
      // device already: the launch goes out with no host round trip, and the
      // front reaches the host by its first synchronisation.
      std::vector<u32> head(pre);
      const std::function<void()> fill = [&]() {
        learned_front_finish(c, order.cand, ms_sched);
        for (u32 i = 0; i < pre; ++i) head[i] = order.next(a_lo);
      };
      r = run_chunk(T, head.data(), pre, 1, &fill, true);
      rs.ms_schedule += ms_sched;
      if (r || T.stop) return r;
      taken.assign(head.begin(), head.end());
      chunk = std::min<u32>(chunk * 4, TCH);
    } else {
      rs.ms_schedule += ms_sched;
    }
    TupleFeed feed(T, 1, chunk);
    while (!order.done()) {
      u32 u = order.next(a_lo);
      taken.push_back(u);
      if ((r = feed.push(&u)) || T.stop) return r;
    }
    std::vector<uint8_t> done(a_hi - a_lo, 0);
    for (u32 u : taken) done[u - a_lo] = 1;
    for (u32 u = a_lo; u < a_hi; ++u)
      if (!done[u - a_lo] && ((r = feed.push(&u)) || T.stop)) return r;
    if ((r = feed.flush()) || T.stop) return r;
    T.slice = 0;
    jfirst = 2;
  }
  // Tuples come in blocks of 4096 first vertices, tuple size outermost within a
  // block.  Plain mode takes the same order one first vertex per launch
  // (per-seed processing); the other modes fill each launch.
  const u32 block = 4096u;
  for (u32 s0 = a_lo; s0 < a_hi && !T.stop; s0 += block) {
    const u32 s1 = std::min(a_hi, s0 + block);
    for (u32 j = jfirst; j <= jmax && !T.stop; ++j) {
      TupleFeed feed(T, j, chunk);
      for (u32 s = s0; s < s1 && !T.stop; ++s) {
        if (j > 1 && s + (j - 1) >= a_hi) continue;
        u32 idx[thin::MAXJ];
        idx[0] = s;
        for (u32 i = 1; i < j; ++i) idx[i] = s + i;
        for (;;) {
          if ((r = feed.push(idx))) return r;
          if (T.stop) break;
          if (j == 1) break;
          int p = (int)j - 1;
          while (p >= 1 && idx[p] == a_hi - (j - p)) --p;
          if (p < 1) break;
          ++idx[p];
          for (u32 q = p + 1; q < j; ++q) idx[q] = idx[q - 1] + 1;
        }
        if (per_seed && (r = feed.flush())) return r;
      }
      if (!T.stop && (r = feed.flush())) return r;
    }
  }
  return GMK_OK;
}

extern "C" int gmk_probe_items(gmk_ctx* c, const gmk_params* prm, uint32_t first, uint32_t count, uint64_t ranks,
                               uint64_t* results, double* lane_ms) {
  if (!c || !prm || !results || !lane_ms || ranks == 0) return GMK_EINVAL;
  if (c->n == 0) { c->err = "no graph loaded"; return GMK_ENOGRAPH; }
  const int side_a = first < c->n_left ? 0 : 1;
  if ((u64)first + count > (side_a ? c->n : c->n_left)) { c->err = "items must lie on one side of the graph"; return GMK_EINVAL; }
  if (prm->k < 1 || prm->k > 3 || prm->theta > 1) {
    c->err = "probes cover one-vertex items: 1 <= k <= 3 and theta <= 1";
    return GMK_EINVAL;
  }
  HIPCHK(c, hipSetDevice(c->device));
  int khz = 0;
  HIPCHK(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
  RunState rs{};
  rs.keep = false;
  rs.limit = ~0ull;
  int r;
  if ((r = ensure(c, c->counters, sizeof(RunCounters)))) return r;
  HIPCHK(c, hipMemsetAsync(c->counters.p, 0, sizeof(RunCounters), c->stream));
  if ((r = ensure(c, c->res, 4))) return r;
  c->res_words = 0;
  c->h_cnt->res_count = 0;
  DeviceVec log;
  if ((r = ensure(c, log, (size_t)TCH * 16))) return r;
  HIPCHK(c, hipMemsetAsync(log.p, 0, (size_t)TCH * 16, c->stream));
  std::vector<u64> res(count, 0), ticks(count, 0);
  ThinRun T;
  gmk_params p2 = *prm;
  p2.mode = GMK_MODE_BATCHED;
  p2.limit = 0;
  const u32 min_b = side_a ? std::max<u32>(std::max<u32>(prm->theta, 1), 2 * prm->k + 1) : std::max<u32>(prm->theta, 1);
  if ((r = thin_setup(c, &p2, rs, side_a, min_b, T))) { (void)hipFree(log.p); return r; }
  T.probe_cap = ranks;
  T.probe_first = first;
  T.probe_res = &res;
  T.probe_ticks = &ticks;
  T.a.item_log = (u64*)log.p;
  u32 chunk = TCH;
  {
    TupleFeed feed(T, 1, chunk);
    for (u32 u = first; u < first + count && !r; ++u) r = feed.push(&u);
    if (!r) r = feed.flush();
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(log.p);
  if (r) return r;
  for (u32 i = 0; i < count; ++i) {
    results[i] = res[i];
    lane_ms[i] = khz > 0 ? (double)ticks[i] / (double)khz : 0.0;
  }
  return GMK_OK;
}

int gmk::build_neighbour_hashes(gmk_ctx* c) {
  c->nh_min = 0xFFFFFFFFu;
  c->nh_slots = 0;
  u32 nh_min = NH_MIN_DEG;
  if (const char* e = getenv("GMK_NHASH_MIN")) nh_min = (u32)std::min<unsigned long>(strtoul(e, nullptr, 10), ~0u);
  if (nh_min == 0) return GMK_OK;
  std::vector<u64> noff(c->n, 0);
  std::vector<u32> verts;  // work items: (vertex, first neighbour of its chunk)
  u64 slots = 0;
  for (u32 v = 0; v < c->n; ++v) {
    const u32 d = c->off[v + 1] - c->off[v];
    if (d < nh_min) continue;
    if (d >= (1u << 30)) { c->err = "a vertex has 2^30 or more neighbours"; return GMK_EUNSUPPORTED; }
    noff[v] = slots;
    slots += 1ull << thin::nh_log2(d);
    for (u32 c0 = 0; c0 < d; c0 += NH_CHUNK) { verts.push_back(v); verts.push_back(c0); }
  }
  if (verts.empty()) return GMK_OK;
  int r;
  if ((r = ensure(c, c->nh_off, (size_t)c->n * 8))) return r;
  if ((r = ensure(c, c->nh, (size_t)slots * 4))) return r;
  if ((r = ensure(c, c->nh_verts, verts.size() * 4))) return r;
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(c->nh_off.p, noff.data(), (size_t)c->n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->nh_verts.p, verts.data(), verts.size() * 4, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemsetAsync(c->nh.p, 0xFF, (size_t)slots * 4, s));
  HIPCHK(c, launch_build_neighbour_hash(c->d_off, c->d_adj, (const u32*)c->nh_verts.p, (u32)(verts.size() / 2),
                                        (const u64*)c->nh_off.p, (u32*)c->nh.p, s));
  HIPCHK(c, hipStreamSynchronize(s));
  c->nh_min = nh_min;
  c->nh_slots = slots;
  return GMK_OK;
}
