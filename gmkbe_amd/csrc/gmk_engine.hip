// Host side of the MI355X maximal k-biplex enumerator: context, graph upload,
// tuning, the single-graph and many-graph entry points of the C ABI declared in
// include/gmk.h (the state-stack engine is in bb_run.hip).
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
      c->n_cu = (u32)prop.multiProcessorCount;
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void**)&c->h_cnt, sizeof(RunCounters), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_ctl, sizeof(LaneCtl) * NLANE, hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess) {
    delete c;
    return GMK_EDEVICE;
  }
  // this device's copy of the static scratch array, and pinned staging for it
  size_t sfl = 0;
  if (gmk::static_scratch(&c->scratch, &sfl) != GMK_OK ||
      hipHostMalloc((void**)&c->h_stage, sfl * sizeof(float), hipHostMallocDefault) != hipSuccess) {
    gmk_destroy(c);
    return GMK_EDEVICE;
  }
  *out = c;
  return GMK_OK;
}

extern "C" int gmk_set_tuning(gmk_ctx* c, const gmk_tuning* t) {
  if (!c || !t) return GMK_EINVAL;
  if (t->version != GMK_TUNING_VERSION) { c->err = "tuning table version mismatch"; return GMK_EINVAL; }
  for (int l = 0; l < 7; ++l) {
    if (t->waves_per_cu[l] > 64 || t->batch_states[l] < 64 || t->batch_states[l] > 65536) {
      c->err = "tuning value out of range";
      return GMK_EINVAL;
    }
  }
  if (t->dedup_load_permille < 100 || t->dedup_load_permille > 900) {
    c->err = "dedup load out of range";
    return GMK_EINVAL;
  }
  for (int q = 0; q < 4; ++q)
    if (t->check_form[q] > 1) { c->err = "tuning value out of range"; return GMK_EINVAL; }
  c->tune = *t;
  return GMK_OK;
}

extern "C" int gmk_get_tuning(gmk_ctx* c, gmk_tuning* t) {
  if (!c || !t) return GMK_EINVAL;
  *t = c->tune;
  return GMK_OK;
}

extern "C" void gmk_destroy(gmk_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (DeviceVec* v : {&c->rows, &c->maps, &c->parts, &c->res, &c->counters, &c->tuples, &c->arena, &c->fams,
                        &c->prefix, &c->tcnt, &c->tdefer, &c->ridx, &c->tcount, &c->nh_off, &c->nh, &c->nh_verts, &c->dedup, &c->gcnt,
                        &c->bcsr_off, &c->bcsr_adj, &c->bcsr_base, &c->rowpids, &c->sched_p, &c->sched_hist,
                        &c->sched_sel, &c->core_deg, &c->core_dead, &c->core_front, &c->core_off, &c->core_adj,
                        &c->wide_scratch, &c->lane_ctl, &c->bfly_cnt, &c->bfly_mark, &c->bfly_deg, &c->bfly_off,
                        &c->bfly_adj, &c->seed_order, &c->seed_meta, &c->seed_scratch, &c->seed_arena, &c->seed_tasks,
                        &c->seed_out})
    if (v->p) (void)hipFree(v->p);
  for (int l = 0; l < NLANE; ++l)
    for (DeviceVec* v : {&c->lane_stack[l], &c->lane_next[l], &c->lane_retry[l], &c->lane_spill[l]})
      if (v->p) (void)hipFree(v->p);
  if (c->d_off) (void)hipFree(c->d_off);
  if (c->d_adj) (void)hipFree(c->d_adj);
  if (c->d_deg) (void)hipFree(c->d_deg);
  if (c->h_cnt) (void)hipHostFree(c->h_cnt);
  if (c->h_ctl) (void)hipHostFree(c->h_ctl);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->ev2) (void)hipEventDestroy(c->ev2);
  if (c->ev3) (void)hipEventDestroy(c->ev3);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char* gmk_last_error(gmk_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int gmk_device_name(gmk_ctx* c, char* out, uint32_t cap) {
  if (!c || !out || cap == 0) return GMK_EINVAL;
  hipDeviceProp_t prop;
  HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
  // the marketing name can be empty where the driver lacks its ID table; the
  // architecture name is always set
  snprintf(out, cap, "%s", prop.name[0] ? prop.name : prop.gcnArchName);
  return GMK_OK;
}

__global__ __launch_bounds__(256) void vertex_degrees(const u32* off, u32 n, u32* deg) {
  const u32 v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) deg[v] = off[v + 1] - off[v];
}

static int upload_graph(gmk_ctx* c) {
  HIPCHK(c, hipSetDevice(c->device));
  if (c->d_off) (void)hipFree(c->d_off);
  if (c->d_adj) (void)hipFree(c->d_adj);
  if (c->d_deg) (void)hipFree(c->d_deg);
  c->d_off = c->d_adj = c->d_deg = nullptr;
  HIPCHK(c, hipMalloc(&c->d_off, sizeof(u32) * (c->n + 1)));
  HIPCHK(c, hipMalloc(&c->d_adj, sizeof(u32) * std::max<size_t>(1, c->adj.size())));
  HIPCHK(c, hipMalloc(&c->d_deg, sizeof(u32) * std::max<size_t>(1, c->n)));
  HIPCHK(c, hipMemcpyAsync(c->d_off, c->off.data(), sizeof(u32) * (c->n + 1), hipMemcpyHostToDevice, c->stream));
  if (!c->adj.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_adj, c->adj.data(), sizeof(u32) * c->adj.size(), hipMemcpyHostToDevice, c->stream));
  if (c->n) hipLaunchKernelGGL(vertex_degrees, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->d_off, c->n, c->d_deg);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return gmk::build_neighbour_hashes(c);
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
  // Symmetry: visiting the vertices in ascending order, the vertices that list
  // w arrive in ascending order, so they must be exactly N(w), in order.  One
  // cursor per vertex walks N(w) as its listers arrive.
  std::vector<u32> cur(c->off.begin(), c->off.end() - 1);
  for (u32 v = 0; v < c->n; ++v)
    for (u32 e = c->off[v]; e < c->off[v + 1]; ++e) {
      const u32 w = c->adj[e];
      if (cur[w] >= c->off[w + 1] || c->adj[cur[w]] != v) {
        c->err = "edge " + std::to_string(v) + " -> " + std::to_string(w) + " has no reverse edge";
        return GMK_EINVAL;
      }
      ++cur[w];
    }
  for (u32 w = 0; w < c->n; ++w)
    if (cur[w] != c->off[w + 1]) {
      c->err = "vertex " + std::to_string(w) + " lists " + std::to_string(c->adj[cur[w]]) +
               ", which does not list it";
      return GMK_EINVAL;
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

void gmk::RunState::report(const char* phase, u64 count, u64 live) {
  if (!verbose) return;
  const double now =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
  if (now - last_print_ms < 10000.0) return;
  last_print_ms = now;
  fprintf(stderr, "[gmk] %.1f s %s: %llu results, %.0f ms on device, %llu launches, %llu live states\n",
          (now - t0_ms) / 1000.0, phase, (unsigned long long)count, ms_expand, (unsigned long long)launches,
          (unsigned long long)live);
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
// Result dedup table growth.
// ---------------------------------------------------------------------------
// One thread per slot of the old table: a filled slot (fingerprint, check word)
// is claimed in the new table at the first empty slot of its probe sequence.
// The old table holds no two slots with both words equal, so no key test is
// needed; every slot's check word was written by the launch that claimed it.
__global__ __launch_bounds__(256) void dedup_rehash(const unsigned long long* old, u64 old_cap,
                                                    unsigned long long* slots, u64 mask) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= old_cap) return;
  const unsigned long long key = old[2 * i];
  if (!key) return;
  u64 h = mix64(key ^ 0xD6E8FEB86659FD93ull) & mask;
  while (atomicCAS(&slots[2 * h], 0ull, key) != 0ull) h = (h + 1) & mask;  // the new table has room: load < 1/2
  slots[2 * h + 1] = old[2 * i + 1];
}

int gmk::dedup_reserve(gmk_ctx* c, RunState& rs, u64 bound) {
  if (!rs.dedup.on) return GMK_OK;
  u64 cap = rs.dedup.mask + 1;
  if ((double)bound <= (double)cap * rs.dedup_load) return GMK_OK;
  u64 ncap = cap;
  while ((double)ncap * rs.dedup_load < (double)bound) ncap <<= 1;
  if (ncap > (1ull << 34)) { c->err = "dedup table would exceed 2^34 slots"; return GMK_EOVERFLOW; }
  DeviceVec nv;
  int r;
  if ((r = ensure(c, nv, ncap * 16))) return r;
  HIPCHK(c, hipMemsetAsync(nv.p, 0, ncap * 16, c->stream));
  const u64 blocks = (cap + 255) / 256;
  hipLaunchKernelGGL(dedup_rehash, dim3((u32)blocks), dim3(256), 0, c->stream, (const unsigned long long*)c->dedup.p,
                     cap, (unsigned long long*)nv.p, ncap - 1);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(c->dedup.p);
  c->dedup = nv;
  rs.dedup.slots = (unsigned long long*)nv.p;
  rs.dedup.mask = ncap - 1;
  return GMK_OK;
}

// ---------------------------------------------------------------------------
// Branch-and-bound over a set of partitions.
// ---------------------------------------------------------------------------
int gmk_pick_w(u32 words) {
  for (int w : {1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024})
    if ((u32)w >= words) return w;
  return 0;
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
  rs.verbose = prm->verbose != 0;
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
  c->h_cnt->res_count = 0;  // read back after every launch; the dedup table is sized from it
  if (prm->dedup) {
    // A limited run sizes the table for the limit; an unlimited one starts at
    // 2^20 slots and doubles (rehashing on the device) whenever the results so
    // far plus the most a launch can add would pass the load factor.
    rs.dedup_load = (prm->dedup_load_permille ? prm->dedup_load_permille : c->tune.dedup_load_permille) / 1000.0;
    const u64 expect = prm->limit ? prm->limit : (1ull << 19);
    u64 cap = 1024;
    while ((double)cap * rs.dedup_load < (double)expect) cap <<= 1;
    if ((r = ensure(c, c->dedup, cap * 16))) return r;
    HIPCHK(c, hipMemsetAsync(c->dedup.p, 0, cap * 16, c->stream));
    rs.dedup.slots = (unsigned long long*)c->dedup.p;
    rs.dedup.mask = cap - 1;
    rs.dedup.stats = ((RunCounters*)c->counters.p)->dstat;
    rs.dedup.on = 1;
  }
  // GMK_DEDUP_REPLAY=1 runs every thin check launch twice and pushes every
  // global root state twice, so each result is produced twice: tests use it to
  // show that the dedup table drops exactly the repeats.
  if (const char* e = getenv("GMK_DEDUP_REPLAY")) rs.replay = atoi(e) != 0;

  const u32 nr = c->n - c->n_left;
  const u32 wl = (c->n_left + 63) / 64, wr = (nr + 63) / 64;
  int part = prm->partition;
  const int Wg = gmk_pick_w(std::max(wl, wr));
  if (part == GMK_PART_AUTO) part = (Wg && Wg <= 16) ? GMK_PART_GLOBAL : GMK_PART_SEEDED;
  if (part == GMK_PART_GLOBAL) {
    if (!Wg) { c->err = "graph too large for a global partition"; return GMK_EUNSUPPORTED; }
    PartitionSet ps;
    build_global_partition(c->n, c->n_left, c->off.data(), c->adj.data(), prm->k, prm->theta, Wg, ps);
    if (rs.replay) ps.roots.insert(ps.roots.end(), ps.roots.begin(), ps.roots.end());
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
    st->dedup_undecided += h.dstat[4 * i + 3];
  }
  if (h.flags & FLAG_TABLE_FULL) { c->err = "dedup table full (raise dedup_load_permille headroom)"; return GMK_EOVERFLOW; }
  st->fp_sum += rs.fp_sum_extra;
  st->fp_xor ^= rs.fp_xor_extra;
  c->res_words = rs.keep ? std::min<u64>(h.res_words, rs.res_cap) : 0;
  st->launches = rs.launches;
  st->spills = rs.spills;
  st->ms_expand = rs.ms_expand;
  st->ms_schedule = rs.ms_schedule;
  st->ms_reduce = rs.ms_reduce;
  st->bytes_reduce = rs.bytes_reduce;
  st->reduce_removed = rs.reduce_removed;
  st->reduce_edges = rs.reduce_edges;
  st->ms_build = rs.ms_build;
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
  rs.verbose = prm->verbose != 0;
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
    // graphs with theta <= 2k run as one global partition, bounded by the widest state
    if (!W[i] && g.theta <= 2 * g.k) {
      c->err = "graph " + std::to_string(i) + " has a side above 65536 vertices and theta <= 2k";
      return GMK_EUNSUPPORTED;
    }
  }
  // schedule
  std::vector<uint32_t> sched(n_graphs);
  for (uint32_t i = 0; i < n_graphs; ++i) sched[i] = order ? order[i] : i;
  if (prm->mode == GMK_MODE_LEARNED) {
    std::vector<double> cost(n_graphs);
    if ((r = gmk_predict_costs(c, graphs, n_graphs, cost.data()))) return r;
    std::stable_sort(sched.begin(), sched.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  }
  // Graphs with theta > 2k run as seed partitions (every result has both sides
  // >= 2k+1, the regime where 2-hop seed partitions are exact); the others are
  // one global partition each.
  std::vector<SeedJob> jobs;
  std::vector<uint32_t> job_graph;
  std::vector<uint32_t> global_graphs;
  for (uint32_t i : sched) {
    const gmk_graph_desc& g = graphs[i];
    if (g.theta > 2 * g.k) {
      jobs.push_back(make_seed_job(HostGraph{g.n, g.n_left, g.off, g.adj}, g.k, g.theta, i));
      job_graph.push_back(i);
    } else {
      global_graphs.push_back(i);
    }
  }
  if (prm->mode == GMK_MODE_PLAIN) {
    for (uint32_t i : global_graphs) {
      const gmk_graph_desc& g = graphs[i];
      PartitionSet ps;
      build_global_partition(g.n, g.n_left, g.off, g.adj, g.k, g.theta, W[i], ps, i);
      const double before = rs.ms_kernel[0];
      if ((r = gmk_run_bb(c, ps, prm, st, rs))) return r;
      per_graph[i].ms = rs.ms_kernel[0] - before;
      ++st->partitions;
    }
    std::vector<double> ms(jobs.size(), 0.0);
    if ((r = run_seed_jobs(c, jobs, prm, st, rs, ms.data()))) return r;
    for (size_t q = 0; q < jobs.size(); ++q) per_graph[job_graph[q]].ms = ms[q];
  } else {
    gmk_params p2 = *prm;
    p2.mode = GMK_MODE_BATCHED;
    // global partitions of every width in one run; the first graph of the
    // schedule is expanded first
    if (!global_graphs.empty()) {
      std::vector<PartitionSet> sets(NLANE);
      size_t rows = 0, maps = 0, parts = 0;
      for (uint32_t i : global_graphs) {
        const gmk_graph_desc& g = graphs[i];
        int l = 0;
        while ((1 << l) < W[i]) ++l;
        build_global_partition(g.n, g.n_left, g.off, g.adj, g.k, g.theta, W[i], sets[l], i);
      }
      for (auto& ps : sets) { rows += ps.rows.size(); maps += ps.maps.size(); parts += ps.parts.size(); }
      BBEngine eng(c, &p2, rs, true);
      if ((r = eng.reserve(rows, maps, parts))) return r;
      for (auto& ps : sets)
        if (!ps.parts.empty() && (r = eng.add_set(ps))) return r;
      if ((r = eng.drain())) return r;
      st->partitions += global_graphs.size();
    }
    if ((r = run_seed_jobs(c, jobs, &p2, st, rs, nullptr))) return r;
    // graphs share the device in this mode; a lone graph's time is the run's
    if (n_graphs == 1) per_graph[0].ms = rs.ms_kernel[0];
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
  st->spills = rs.spills;
  st->ms_expand = rs.ms_expand;
  st->ms_build = rs.ms_build;
  st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->prog_count = std::move(rs.prog_count);
  c->prog_ms = std::move(rs.prog_ms);
  return GMK_OK;
}
