// Edge-list parsing for the .graph.bin converter (gmkbe_amd/convert.py), on
// host threads: the role of graph_convert's parallel parse
// (preprocess/graph_conver.cpp:505-552), with the same rules.  The file is cut
// into one byte range per thread; a range starts after a newline (its first
// line belongs to the previous range) and ends after one.  In a range, blanks
// and line ends are skipped, a line starting with '%' or '#' is a comment, and
// otherwise two numbers are read, each skipping any non-digits before it (so a
// line with one number takes the next line's first, within the range), and the
// rest of the line is dropped.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/gmk.h"

namespace {

inline uint32_t read_uint(const char*& p, const char* pend) {
  while (p < pend && (*p < '0' || *p > '9')) ++p;
  uint32_t val = 0;
  while (p < pend && *p >= '0' && *p <= '9') {
    val = val * 10u + (uint32_t)(*p - '0');
    ++p;
  }
  return val;
}

void parse_range(const char* p, const char* pend, std::vector<uint32_t>& us, std::vector<uint32_t>& vs) {
  while (p < pend) {
    const char ch = *p;
    if (ch == '\n' || ch == '\r' || ch == ' ' || ch == '\t') { ++p; continue; }
    if (ch == '%' || ch == '#') {
      while (p < pend && *p != '\n') ++p;
      continue;
    }
    const uint32_t u = read_uint(p, pend);
    const uint32_t v = read_uint(p, pend);
    us.push_back(u);
    vs.push_back(v);
    while (p < pend && *p != '\n') ++p;
  }
}

}  // namespace

extern "C" int gmk_parse_edge_list(const char* path, int threads, uint32_t** u_out, uint32_t** v_out,
                                   uint64_t* m_out) {
  if (!path || !u_out || !v_out || !m_out) return GMK_EINVAL;
  *u_out = *v_out = nullptr;
  *m_out = 0;
  FILE* f = fopen(path, "rb");
  if (!f) return GMK_EIO;
  std::vector<char> data;
  if (fseek(f, 0, SEEK_END) == 0) {
    const long sz = ftell(f);
    if (sz > 0) {
      data.resize((size_t)sz);
      fseek(f, 0, SEEK_SET);
      if (fread(data.data(), 1, data.size(), f) != data.size()) { fclose(f); return GMK_EIO; }
    }
  }
  fclose(f);
  const size_t fsize = data.size();
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  const int nt = threads;
  std::vector<std::vector<uint32_t>> us(nt), vs(nt);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t]() {
      const char* d = data.data();
      const size_t chunk = fsize / nt;
      size_t beg = (size_t)t * chunk, end = t == nt - 1 ? fsize : beg + chunk;
      if (t != 0 && beg > 0)
        while (beg < fsize && d[beg - 1] != '\n') ++beg;
      if (t != nt - 1 && end > 0)
        while (end < fsize && d[end - 1] != '\n') ++end;
      if (beg < end) {
        us[t].reserve((end - beg) / 12 + 16);
        vs[t].reserve((end - beg) / 12 + 16);
        parse_range(d + beg, d + end, us[t], vs[t]);
      }
    });
  for (auto& th : pool) th.join();
  uint64_t m = 0;
  for (int t = 0; t < nt; ++t) m += us[t].size();
  uint32_t* u = (uint32_t*)malloc(sizeof(uint32_t) * (m + 1));
  uint32_t* v = (uint32_t*)malloc(sizeof(uint32_t) * (m + 1));
  if (!u || !v) { free(u); free(v); return GMK_ENOMEM; }
  uint64_t at = 0;
  for (int t = 0; t < nt; ++t) {
    if (!us[t].empty()) {
      memcpy(u + at, us[t].data(), us[t].size() * 4);
      memcpy(v + at, vs[t].data(), vs[t].size() * 4);
    }
    at += us[t].size();
  }
  *u_out = u;
  *v_out = v;
  *m_out = m;
  return GMK_OK;
}

extern "C" void gmk_free(void* p) { free(p); }
