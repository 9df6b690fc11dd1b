// Host orchestration of the thin-result kernels: tuples are generated seed by
// seed (ascending IDs on the tuple side), families built on the device in
// chunks, then every concrete combination is checked on the device, rank range
// by rank range (each range: rank index, first pass, deferred pass), until the
// result limit is reached.  Also builds the graph's neighbour hashes at upload.
#include <algorithm>
#include <cstddef>
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
};

constexpr u32 TCH = 1u << 16;          // tuples per family-construction launch
constexpr u64 ARENA = 1ull << 27;      // arena words (512 MiB)
constexpr u32 FCAP = 1u << 18;         // families per launch
constexpr u64 RCH = 1ull << 22;        // ranks per check launch
constexpr u32 DEFER_BUDGET = 64;       // candidates one lane tests before a rank goes to a whole wave

struct ThinRun {
  gmk_ctx* c;
  const gmk_params* prm;
  RunState* rs;
  ThinArgs a{};
  std::vector<u64> totals, prefix;
  bool stop = false;
};

int run_chunk(ThinRun& T, const u32* tuples, u32 count, u32 j) {
  gmk_ctx* c = T.c;
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(c->tuples.p, tuples, (size_t)count * j * 4, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemsetAsync(c->tcnt.p, 0, sizeof(ThinCounters), s));
  ThinArgs& a = T.a;
  a.ntuples = count;
  a.j = j;
  HIPCHK(c, hipEventRecord(c->ev0, s));
  HIPCHK(c, launch_thin_build(a, s));
  HIPCHK(c, hipEventRecord(c->ev1, s));
  ThinCounters tc;
  HIPCHK(c, hipMemcpyAsync(&tc, c->tcnt.p, sizeof(tc), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
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
  HIPCHK(c, launch_thin_scan((const Family*)c->fams.p, nfam, (u64*)c->prefix.p, s));
  u64 total = 0;
  HIPCHK(c, hipMemcpyAsync(&total, (u64*)c->prefix.p + nfam, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  a.nfam = nfam;
  RunCounters* dc = (RunCounters*)c->counters.p;
  // GMK_DEDUP_REPLAY: every rank range a second time (its results repeat)
  for (u64 q = 0, nq = T.rs->replay ? 2 : 1; q < nq * ((total + RCH - 1) / RCH); ++q) {
    const u64 r0 = (q % ((total + RCH - 1) / RCH)) * RCH;
    a.r0 = r0;
    a.r1 = std::min(total, r0 + RCH);
    int rr;
    if ((rr = dedup_reserve(c, *T.rs, c->h_cnt->res_count + (a.r1 - a.r0)))) return rr;
    a.dedup = T.rs->dedup;
    // work, defer_n and dwork are consecutive
    HIPCHK(c, hipMemsetAsync(&((ThinCounters*)c->tcnt.p)->work, 0, 3 * sizeof(u32), s));
    HIPCHK(c, hipEventRecord(c->ev0, s));
    HIPCHK(c, launch_thin_rank_index((const u64*)c->prefix.p, nfam, a.r0, a.r1, (u32*)c->ridx.p, s));
    HIPCHK(c, launch_thin_check(a, s, c->n_cu * 8, c->tune.check_form[a.k < 4 ? a.k : 3]));
    HIPCHK(c, launch_thin_check_deferred(a, s, c->n_cu * 4));
    HIPCHK(c, hipEventRecord(c->ev1, s));
    HIPCHK(c, hipMemcpyAsync(c->h_cnt, dc, offsetof(RunCounters, fp), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    T.rs->ms_expand += ms;
    T.rs->ms_kernel[2] += ms;
    T.rs->launches_kernel[2]++;
    T.rs->launches++;
    T.rs->progress(std::min<u64>(c->h_cnt->res_count, T.rs->limit));
    T.rs->report(a.swap_sides ? "thin (tuples on the right)" : "thin (tuples on the left)",
                 std::min<u64>(c->h_cnt->res_count, T.rs->limit), 0);
    if (c->h_cnt->flags & FLAG_RES_OVERFLOW) { c->err = "result buffer overflow (raise result_capacity_words)"; return GMK_EOVERFLOW; }
    if (c->h_cnt->flags & FLAG_LIMIT) { T.stop = true; T.rs->hit_limit = true; return GMK_OK; }
  }
  return GMK_OK;
}

}  // namespace

int gmk_run_thin(gmk_ctx* c, const gmk_params* prm, gmk_stats* st, RunState& rs, int side_a, uint32_t min_b) {
  (void)st;
  const u32 k = prm->k;
  const u32 jmin = std::max<u32>(prm->theta, 1);
  const u32 jmax = std::min<u32>(2 * k, thin::MAXJ);
  if (jmin > jmax) return GMK_OK;
  int r;
  if ((r = ensure(c, c->tuples, (size_t)TCH * thin::MAXJ * 4))) return r;
  if ((r = ensure(c, c->arena, ARENA * 4))) return r;
  if ((r = ensure(c, c->fams, (size_t)FCAP * sizeof(Family)))) return r;
  if ((r = ensure(c, c->prefix, (size_t)(FCAP + 1) * 8))) return r;
  if ((r = ensure(c, c->tcnt, sizeof(ThinCounters)))) return r;
  if ((r = ensure(c, c->tdefer, RCH * 8))) return r;
  if ((r = ensure(c, c->ridx, ((RCH >> RIDX_LOG) + 2) * 4))) return r;
  if ((r = ensure(c, c->tcount, (size_t)TCH * 8 * 4))) return r;
  ThinRun T;
  T.c = c; T.prm = prm; T.rs = &rs;
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

  const u32 a_lo = a.g.a_lo, a_hi = a.g.a_hi;
  const u32* off = c->off.data();
  // Tuples come in blocks of 4096 first vertices, tuple size outermost within a
  // block.  Plain mode takes the same order one first vertex per launch
  // (per-seed processing); the other modes fill each launch.
  const u32 block = 4096u;
  const bool per_seed = prm->mode == GMK_MODE_PLAIN;
  // Launch sizes ramp up from 64 tuples (x4 per launch) to TCH, so a run that
  // reaches its limit within the first tuples stops early; the order is the same.
  u32 chunk = 64;
  std::vector<u32> buf;
  buf.reserve((size_t)TCH * thin::MAXJ);
  for (u32 s0 = a_lo; s0 < a_hi && !T.stop; s0 += block) {
    const u32 s1 = std::min(a_hi, s0 + block);
    for (u32 j = jmin; j <= jmax && !T.stop; ++j) {
      buf.clear();
      u32 count = 0;
      u64 est = 0;
      auto flush = [&]() -> int {
        if (count == 0) return GMK_OK;
        int rr = run_chunk(T, buf.data(), count, j);
        buf.clear(); count = 0; est = 0;
        chunk = std::min<u32>(chunk * 4, TCH);
        return rr;
      };
      for (u32 s = s0; s < s1 && !T.stop; ++s) {
        if (j > 1 && s + (j - 1) >= a_hi) continue;
        u32 idx[thin::MAXJ];
        idx[0] = s;
        for (u32 i = 1; i < j; ++i) idx[i] = s + i;
        for (;;) {
          u64 need = 0;
          for (u32 i = 0; i < j; ++i) need += 2ull * (off[idx[i] + 1] - off[idx[i]]);
          if (count >= chunk || est + need > ARENA) { if ((r = flush())) return r; if (T.stop) break; }
          buf.insert(buf.end(), idx, idx + j);
          ++count;
          est += need;
          if (j == 1) break;
          int p = (int)j - 1;
          while (p >= 1 && idx[p] == a_hi - (j - p)) --p;
          if (p < 1) break;
          ++idx[p];
          for (u32 q = p + 1; q < j; ++q) idx[q] = idx[q - 1] + 1;
        }
        if (per_seed && (r = flush())) return r;
      }
      if (!T.stop && (r = flush())) return r;
    }
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
