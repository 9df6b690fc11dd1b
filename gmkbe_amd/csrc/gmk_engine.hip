// Host side of the MI355X maximal k-biplex enumerator: context, graph upload,
// partition construction, the batch loop that drains the device state stack, and
// the C ABI declared in include/gmk.h.
#include "gmk_device.h"
#include "gmk_host.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gmk.h"

using namespace gmk;

namespace gmk {
hipError_t launch_bb_expand(int W, u32 n_states, hipStream_t s, const ExpandArgs& a, u32 persistent_waves);
}

extern "C" int gmk_create(int device, gmk_ctx** out) {
  if (!out) return GMK_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return GMK_EDEVICE;
  if (device < 0 || device >= ndev) return GMK_EINVAL;
  gmk_ctx* c = new gmk_ctx();
  c->device = device;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      c->persistent_waves = (u32)prop.multiProcessorCount * 16;
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void**)&c->h_cnt, sizeof(RunCounters), hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return GMK_EDEVICE;
  }
  *out = c;
  return GMK_OK;
}

extern "C" void gmk_destroy(gmk_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (DeviceVec* v : {&c->stack, &c->next, &c->retry, &c->rows, &c->maps, &c->parts, &c->res, &c->counters,
                        &c->tuples, &c->arena, &c->fams, &c->prefix, &c->tcnt, &c->dedup, &c->gcnt})
    if (v->p) (void)hipFree(v->p);
  if (c->d_off) (void)hipFree(c->d_off);
  if (c->d_adj) (void)hipFree(c->d_adj);
  if (c->h_cnt) (void)hipHostFree(c->h_cnt);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char* gmk_last_error(gmk_ctx* c) { return c ? c->err.c_str() : "null context"; }

static int upload_graph(gmk_ctx* c) {
  HIPCHK(c, hipSetDevice(c->device));
  if (c->d_off) (void)hipFree(c->d_off);
  if (c->d_adj) (void)hipFree(c->d_adj);
  c->d_off = c->d_adj = nullptr;
  HIPCHK(c, hipMalloc(&c->d_off, sizeof(u32) * (c->n + 1)));
  HIPCHK(c, hipMalloc(&c->d_adj, sizeof(u32) * std::max<size_t>(1, c->adj.size())));
  HIPCHK(c, hipMemcpyAsync(c->d_off, c->off.data(), sizeof(u32) * (c->n + 1), hipMemcpyHostToDevice, c->stream));
  if (!c->adj.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_adj, c->adj.data(), sizeof(u32) * c->adj.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GMK_OK;
}

static int check_csr(gmk_ctx* c) {
  if (c->n_left > c->n) { c->err = "n_left > n"; return GMK_EINVAL; }
  if (c->off.size() != (size_t)c->n + 1 || c->off[0] != 0) { c->err = "bad offsets"; return GMK_EINVAL; }
  for (u32 v = 0; v < c->n; ++v) {
    if (c->off[v + 1] < c->off[v] || c->off[v + 1] > c->adj.size()) { c->err = "offsets not monotone"; return GMK_EINVAL; }
    const bool left = v < c->n_left;
    for (u32 e = c->off[v]; e < c->off[v + 1]; ++e) {
      const u32 w = c->adj[e];
      if (w >= c->n || (w < c->n_left) == left) {
        c->err = "neighbour " + std::to_string(w) + " of vertex " + std::to_string(v) + " is not on the other side";
        return GMK_EINVAL;
      }
      if (e > c->off[v] && c->adj[e - 1] >= w) {
        c->err = "neighbour list of vertex " + std::to_string(v) + " is not strictly ascending";
        return GMK_EINVAL;
      }
    }
  }
  return GMK_OK;
}

extern "C" int gmk_set_graph_csr(gmk_ctx* c, uint32_t n, uint32_t n_left, const uint32_t* off, const uint32_t* adj) {
  if (!c || !off || (!adj && off[n] > 0)) return GMK_EINVAL;
  c->n = n; c->n_left = n_left;
  c->order_valid = false;
  c->off.assign(off, off + n + 1);
  c->adj.assign(adj, adj + off[n]);
  int r = check_csr(c);
  if (r != GMK_OK) { c->n = 0; return r; }
  return upload_graph(c);
}

// .graph.bin reader (layout of preprocess/graph_conver.cpp:787-809, read as in
// data/src/tmp_load.h:213-295, with the size check of SGMD22-MBPE/include/Util.h:212-218).
extern "C" int gmk_load_graph_bin(gmk_ctx* c, const char* path, uint32_t* n_out, uint32_t* nl_out, uint64_t* nnz_out) {
  if (!c || !path) return GMK_EINVAL;
  FILE* f = fopen(path, "rb");
  if (!f) { c->err = std::string("cannot open ") + path; return GMK_EIO; }
  u32 hdr[3];
  if (fread(hdr, 4, 3, f) != 3) { fclose(f); c->err = "truncated header"; return GMK_EIO; }
  const u32 n_edge = hdr[0], n = hdr[1], nl = hdr[2];
  fseek(f, 0, SEEK_END);
  const long long fsz = ftell(f);
  fseek(f, 12, SEEK_SET);
  if (nl > n || (u64)fsz != 12ull + 8ull * n + 4ull * n_edge) {
    fclose(f);
    c->err = std::string(path) + ": file size does not match header";
    return GMK_EIO;
  }
  std::vector<u32> off(n + 1, 0), adj(n_edge);
  u64 w = 0;
  for (u32 i = 0; i < n; ++i) {
    u32 rec[2];
    if (fread(rec, 4, 2, f) != 2 || rec[0] != i || w + rec[1] > n_edge) {
      fclose(f);
      c->err = std::string(path) + ": bad vertex record " + std::to_string(i);
      return GMK_EIO;
    }
    if (rec[1] && fread(adj.data() + w, 4, rec[1], f) != rec[1]) {
      fclose(f);
      c->err = std::string(path) + ": truncated neighbour list";
      return GMK_EIO;
    }
    w += rec[1];
    off[i + 1] = (u32)w;
  }
  fclose(f);
  int r = gmk_set_graph_csr(c, n, nl, off.data(), adj.data());
  if (r != GMK_OK) return r;
  if (n_out) *n_out = n;
  if (nl_out) *nl_out = nl;
  if (nnz_out) *nnz_out = n_edge;
  return GMK_OK;
}

void gmk::RunState::progress(u64 count) {
  const double now =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
  prog_count.push_back(count);
  prog_ms.push_back(now - t0_ms);
}

extern "C" int gmk_fetch_progress(gmk_ctx* c, uint64_t* counts, double* ms, uint32_t cap, uint32_t* n) {
  if (!c) return GMK_EINVAL;
  const uint32_t len = (uint32_t)c->prog_count.size();
  if (n) *n = len;
  for (uint32_t i = 0; i < len && i < cap; ++i) {
    if (counts) counts[i] = c->prog_count[i];
    if (ms) ms[i] = c->prog_ms[i];
  }
  return GMK_OK;
}

extern "C" int gmk_fetch_results(gmk_ctx* c, uint32_t* words, uint64_t cap, uint64_t* nwords) {
  if (!c) return GMK_EINVAL;
  if (nwords) *nwords = c->res_words;
  if (!words) return GMK_OK;
  if (cap < c->res_words) return GMK_EOVERFLOW;
  if (c->res_words) {
    HIPCHK(c, hipMemcpyAsync(words, c->res.p, sizeof(u32) * c->res_words, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return GMK_OK;
}

// ---------------------------------------------------------------------------
// Branch-and-bound over a set of partitions.
// ---------------------------------------------------------------------------
int gmk_pick_w(u32 words) {
  for (int w : {1, 2, 4, 8, 16, 32, 64})
    if ((u32)w >= words) return w;
  return 0;
}

int gmk_run_bb(gmk_ctx* c, const PartitionSet& ps, const gmk_params* prm, gmk_stats* st, RunState& rs) {
  const int W = ps.W;
  const u64 stride = 1 + 6ull * W;
  const u64 state_bytes = stride * 8;
  // Device stack: ~1.5 GiB, children buffer half of it.
  const u64 cap = std::max<u64>(1 << 16, (3ull << 29) / state_bytes);
  const u64 ncap = cap / 2;
  const u32 batch_max = 1u << 16;
  int r;
  if ((r = ensure(c, c->stack, cap * state_bytes))) return r;
  if ((r = ensure(c, c->next, ncap * state_bytes))) return r;
  if ((r = ensure(c, c->retry, (u64)batch_max * state_bytes))) return r;
  if ((r = ensure(c, c->rows, std::max<size_t>(8, ps.rows.size() * 8)))) return r;
  if ((r = ensure(c, c->maps, std::max<size_t>(4, ps.maps.size() * 4)))) return r;
  if ((r = ensure(c, c->parts, std::max<size_t>(sizeof(PartDesc), ps.parts.size() * sizeof(PartDesc))))) return r;
  if (ps.roots.size() / stride > cap) { c->err = "too many root states"; return GMK_EOVERFLOW; }
  hipStream_t s = c->stream;
  HIPCHK(c, hipMemcpyAsync(c->rows.p, ps.rows.data(), ps.rows.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->maps.p, ps.maps.data(), ps.maps.size() * 4, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->parts.p, ps.parts.data(), ps.parts.size() * sizeof(PartDesc), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->stack.p, ps.roots.data(), ps.roots.size() * 8, hipMemcpyHostToDevice, s));
  u64 top = ps.roots.size() / stride;

  RunCounters* dc = (RunCounters*)c->counters.p;
  ExpandArgs a{};
  a.parts = (const PartDesc*)c->parts.p;
  a.rows = (const u64*)c->rows.p;
  a.maps = (const u32*)c->maps.p;
  a.k = prm->k;
  a.keep = rs.keep;
  a.res = (u32*)c->res.p;
  a.res_cap = rs.res_cap;
  a.res_words = &dc->res_words;
  a.res_count = &dc->res_count;
  a.limit = rs.limit;
  a.fp = dc->fp;
  a.stat = dc->stat;
  a.flags = &dc->flags;
  a.dedup = rs.dedup;
  a.gcount = rs.gcount;
  a.gfp = rs.gfp;
  a.gcap = rs.gcap;
  a.out_n = &dc->out_n;
  a.retry_n = &dc->retry_n;
  a.work_next = &dc->pad;
  a.retry = (u64*)c->retry.p;
  a.out = (u64*)c->next.p;
  u64* stack = (u64*)c->stack.p;
  const u32 max_children = 2 * 64 * (u32)W + 2;

  while (top > 0) {
    const u32 B = (u32)std::min<u64>(top, batch_max);
    const u64 room = cap - top;
    a.out_cap = (u32)std::min<u64>(ncap, room);
    if (a.out_cap < max_children) { c->err = "state stack full"; return GMK_EOVERFLOW; }
    a.in = stack + (top - B) * stride;
    a.n_in = B;
    HIPCHK(c, hipMemsetAsync(dc, 0, 2 * sizeof(u32), s));
    HIPCHK(c, hipMemsetAsync(&dc->pad, 0, sizeof(u32), s));
    HIPCHK(c, hipEventRecord(c->ev0, s));
    HIPCHK(c, launch_bb_expand(W, B, s, a, prm->mode == GMK_MODE_PLAIN ? 0u : c->persistent_waves));
    HIPCHK(c, hipEventRecord(c->ev1, s));
    HIPCHK(c, hipMemcpyAsync(c->h_cnt, dc, offsetof(RunCounters, fp), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    rs.ms_expand += ms;
    ++rs.launches;
    rs.ms_kernel[0] += ms;
    ++rs.launches_kernel[0];
    const RunCounters& h = *c->h_cnt;
    rs.progress(std::min<u64>(h.res_count, rs.limit));
    if (h.flags & FLAG_RES_OVERFLOW) {
      c->err = "result buffer overflow (raise result_capacity_words)";
      return GMK_EOVERFLOW;
    }
    const u64 nout = std::min<u64>(h.out_n, a.out_cap);
    const u64 nret = h.retry_n;
    u64 dst = top - B;
    if (nret) HIPCHK(c, hipMemcpyAsync(stack + dst * stride, a.retry, nret * state_bytes, hipMemcpyDeviceToDevice, s));
    dst += nret;
    if (nout) HIPCHK(c, hipMemcpyAsync(stack + dst * stride, a.out, nout * state_bytes, hipMemcpyDeviceToDevice, s));
    top = dst + nout;
    if (h.flags & FLAG_LIMIT) { rs.hit_limit = true; break; }
    if (nout == 0 && nret == B && B > 0 && top > 0 && h.out_n >= a.out_cap) {
      // every parent was requeued: only possible when the buffer is smaller than
      // one state's children, which the room check above excludes
      c->err = "no progress in state expansion";
      return GMK_EOVERFLOW;
    }
  }
  return GMK_OK;
}

extern "C" int gmk_enumerate(gmk_ctx* c, const gmk_params* prm, gmk_stats* st) {
  if (!c || !prm) return GMK_EINVAL;
  if (c->n == 0) { c->err = "no graph loaded"; return GMK_ENOGRAPH; }
  HIPCHK(c, hipSetDevice(c->device));
  gmk_stats local{};
  if (!st) st = &local;
  memset(st, 0, sizeof(*st));
  const auto t0 = std::chrono::steady_clock::now();
  RunState rs{};
  rs.t0_ms = std::chrono::duration<double, std::milli>(t0.time_since_epoch()).count();
  rs.keep = prm->keep_results != 0;
  rs.limit = prm->limit ? prm->limit : ~0ull;
  int r;
  if ((r = ensure(c, c->counters, sizeof(RunCounters)))) return r;
  HIPCHK(c, hipMemsetAsync(c->counters.p, 0, sizeof(RunCounters), c->stream));
  if (rs.keep) {
    u64 capw = prm->result_capacity_words;
    if (!capw) capw = prm->limit ? std::min<u64>(1ull << 31, std::max<u64>(1 << 20, prm->limit * 64)) : (1ull << 28);
    if ((r = ensure(c, c->res, capw * 4))) return r;
    rs.res_cap = capw;
  } else {
    if ((r = ensure(c, c->res, 4))) return r;
    rs.res_cap = 0;
  }
  c->res_words = 0;
  if (prm->dedup) {
    const u64 expect = prm->limit ? prm->limit : (1ull << 25);
    const double lf = (prm->dedup_load_permille ? prm->dedup_load_permille : 500) / 1000.0;
    u64 cap = 1024;
    while ((double)cap * lf < (double)expect) cap <<= 1;
    if ((r = ensure(c, c->dedup, cap * 8))) return r;
    HIPCHK(c, hipMemsetAsync(c->dedup.p, 0, cap * 8, c->stream));
    rs.dedup.slots = (unsigned long long*)c->dedup.p;
    rs.dedup.mask = cap - 1;
    rs.dedup.stats = ((RunCounters*)c->counters.p)->dstat;
    rs.dedup.on = 1;
  }

  const u32 nr = c->n - c->n_left;
  const u32 wl = (c->n_left + 63) / 64, wr = (nr + 63) / 64;
  int part = prm->partition;
  const int Wg = gmk_pick_w(std::max(wl, wr));
  if (part == GMK_PART_AUTO) part = (Wg && Wg <= 16) ? GMK_PART_GLOBAL : GMK_PART_SEEDED;
  if (part == GMK_PART_GLOBAL) {
    if (!Wg) { c->err = "graph too large for a global partition"; return GMK_EUNSUPPORTED; }
    PartitionSet ps;
    build_global_partition(c->n, c->n_left, c->off.data(), c->adj.data(), prm->k, prm->theta, Wg, ps);
    st->partitions = 1;
    if ((r = gmk_run_bb(c, ps, prm, st, rs))) return r;
  } else {
    if ((r = gmk_run_seeded(c, prm, st, rs))) return r;
  }

  // Read back totals.
  HIPCHK(c, hipMemcpyAsync(c->h_cnt, c->counters.p, sizeof(RunCounters), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const RunCounters& h = *c->h_cnt;
  st->count = std::min<u64>(h.res_count, rs.limit);
  for (int i = 0; i < NSTRIPE; ++i) {
    st->fp_sum += h.fp[2 * i];
    st->fp_xor ^= h.fp[2 * i + 1];
    st->states += h.stat[NSTAT * i];
    st->intersections += h.stat[NSTAT * i + 1];
    for (int q = 0; q < 4; ++q) st->bytes_kernel[q] += h.stat[NSTAT * i + 3 + q];
  }
  for (int q = 0; q < 4; ++q) {
    st->algo_bytes += st->bytes_kernel[q];
    st->ms_kernel[q] = rs.ms_kernel[q];
    st->launches_kernel[q] = rs.launches_kernel[q];
  }
  for (int i = 0; i < NSTRIPE; ++i) {
    st->dedup_inserts += h.dstat[4 * i];
    st->dedup_probes += h.dstat[4 * i + 1];
    st->dedup_dups += h.dstat[4 * i + 2];
  }
  if (h.flags & FLAG_TABLE_FULL) { c->err = "dedup table full (raise dedup_load_permille headroom)"; return GMK_EOVERFLOW; }
  st->fp_sum += rs.fp_sum_extra;
  st->fp_xor ^= rs.fp_xor_extra;
  c->res_words = rs.keep ? std::min<u64>(h.res_words, rs.res_cap) : 0;
  st->launches = rs.launches;
  st->ms_expand = rs.ms_expand;
  st->hit_limit = rs.hit_limit || h.res_count >= rs.limit;
  st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  rs.progress(st->count);
  c->prog_count = std::move(rs.prog_count);
  c->prog_ms = std::move(rs.prog_ms);
  return GMK_OK;
}

// ---------------------------------------------------------------------------
// Many graphs in one run.
// ---------------------------------------------------------------------------
extern "C" int gmk_enumerate_batch(gmk_ctx* c, const gmk_graph_desc* graphs, uint32_t n_graphs,
                                   const gmk_params* prm, const uint32_t* order, gmk_graph_stats* per_graph,
                                   gmk_stats* st) {
  if (!c || !graphs || !prm || !per_graph) return GMK_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  gmk_stats local{};
  if (!st) st = &local;
  memset(st, 0, sizeof(*st));
  memset(per_graph, 0, sizeof(gmk_graph_stats) * n_graphs);
  const auto t0 = std::chrono::steady_clock::now();
  int r;
  if ((r = ensure(c, c->counters, sizeof(RunCounters)))) return r;
  HIPCHK(c, hipMemsetAsync(c->counters.p, 0, sizeof(RunCounters), c->stream));
  if ((r = ensure(c, c->gcnt, (size_t)n_graphs * 24 + 24))) return r;
  HIPCHK(c, hipMemsetAsync(c->gcnt.p, 0, (size_t)n_graphs * 24 + 24, c->stream));
  if ((r = ensure(c, c->res, 4))) return r;
  RunState rs{};
  rs.t0_ms = std::chrono::duration<double, std::milli>(t0.time_since_epoch()).count();
  rs.keep = false;
  rs.limit = ~0ull;
  rs.gcount = (unsigned long long*)c->gcnt.p;
  rs.gfp = (u64*)((char*)c->gcnt.p + (size_t)n_graphs * 8);
  rs.gcap = prm->limit ? prm->limit : ~0ull;
  c->res_words = 0;
  // widths and validation
  std::vector<int> W(n_graphs);
  for (uint32_t i = 0; i < n_graphs; ++i) {
    const gmk_graph_desc& g = graphs[i];
    if (!g.off || (!g.adj && g.off[g.n] > 0) || g.n_left > g.n || g.theta == 0) { c->err = "bad graph " + std::to_string(i); return GMK_EINVAL; }
    const u32 nr = g.n - g.n_left;
    W[i] = gmk_pick_w(std::max((g.n_left + 63) / 64, (nr + 63) / 64));
    if (!W[i]) { c->err = "graph " + std::to_string(i) + " has a side above 4096 vertices"; return GMK_EUNSUPPORTED; }
  }
  // schedule
  std::vector<uint32_t> sched(n_graphs);
  for (uint32_t i = 0; i < n_graphs; ++i) sched[i] = order ? order[i] : i;
  if (prm->mode == GMK_MODE_LEARNED) {
    std::vector<double> cost(n_graphs);
    if ((r = gmk_predict_costs(c, graphs, n_graphs, cost.data()))) return r;
    std::stable_sort(sched.begin(), sched.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  }
  if (prm->mode == GMK_MODE_PLAIN) {
    for (uint32_t i : sched) {
      const gmk_graph_desc& g = graphs[i];
      PartitionSet ps;
      build_global_partition(g.n, g.n_left, g.off, g.adj, g.k, g.theta, W[i], ps, i);
      const double before = rs.ms_kernel[0];
      if ((r = gmk_run_bb(c, ps, prm, st, rs))) return r;
      per_graph[i].ms = rs.ms_kernel[0] - before;
      ++st->partitions;
    }
  } else {
    // one run per width; graphs pushed so that the first in the schedule ends
    // on top of the stack (taken first)
    for (int w : {1, 2, 4, 8, 16, 32, 64}) {
      PartitionSet ps;
      u32 cnt = 0;
      for (auto it = sched.rbegin(); it != sched.rend(); ++it) {
        const uint32_t i = *it;
        if (W[i] != w) continue;
        const gmk_graph_desc& g = graphs[i];
        build_global_partition(g.n, g.n_left, g.off, g.adj, g.k, g.theta, w, ps, i);
        ++cnt;
      }
      if (!cnt) continue;
      gmk_params p2 = *prm;
      p2.mode = GMK_MODE_BATCHED;
      if ((r = gmk_run_bb(c, ps, &p2, st, rs))) return r;
      st->partitions += cnt;
    }
  }
  HIPCHK(c, hipMemcpyAsync(c->h_cnt, c->counters.p, sizeof(RunCounters), hipMemcpyDeviceToHost, c->stream));
  std::vector<u64> gc((size_t)n_graphs * 3);
  HIPCHK(c, hipMemcpyAsync(gc.data(), c->gcnt.p, (size_t)n_graphs * 24, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const RunCounters& h = *c->h_cnt;
  for (uint32_t i = 0; i < n_graphs; ++i) {
    per_graph[i].count = std::min<u64>(gc[i], rs.gcap);
    per_graph[i].hit_cap = gc[i] >= rs.gcap;
    per_graph[i].fp_sum = gc[n_graphs + 2 * i];
    per_graph[i].fp_xor = gc[n_graphs + 2 * i + 1];
    st->count += per_graph[i].count;
    st->fp_sum += per_graph[i].fp_sum;
    st->fp_xor ^= per_graph[i].fp_xor;
  }
  for (int i = 0; i < NSTRIPE; ++i) {
    st->states += h.stat[NSTAT * i];
    st->intersections += h.stat[NSTAT * i + 1];
    for (int q = 0; q < 4; ++q) st->bytes_kernel[q] += h.stat[NSTAT * i + 3 + q];
  }
  for (int q = 0; q < 4; ++q) {
    st->algo_bytes += st->bytes_kernel[q];
    st->ms_kernel[q] = rs.ms_kernel[q];
    st->launches_kernel[q] = rs.launches_kernel[q];
  }
  st->launches = rs.launches;
  st->ms_expand = rs.ms_expand;
  st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->prog_count = std::move(rs.prog_count);
  c->prog_ms = std::move(rs.prog_ms);
  return GMK_OK;
}
