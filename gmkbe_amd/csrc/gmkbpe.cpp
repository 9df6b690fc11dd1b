// gmkbpe: command-line driver over libgmk.so (include/gmk.h).
//
// Flags and defaults follow the reference driver (src/main.cu:1719-1775):
//   -f <file.graph.bin>  graph
//   -k <k>               default 1
//   -r <limit>           default 10000000 (0 = no limit)
//   -d <device>          default 0
//   -q <0|1>             print "Result n: X: ... | Y: ..." lines (default 0)
//   -m <0|1>             verbose (per-kernel statistics)
//   -g <n>               accepted and ignored, as in the reference
// Additions: -t <theta> (size lower bound, default 1), --mode plain|batched|learned,
// --partition auto|global|seeded, --tune <table.json> (launch table written by
// gmkbe_amd/autotune.py; the library's built-in MI355X table otherwise),
// --format gmkbpe|bppivot|itraversal (the -q 1 result lines of the reference,
// of its BPPivot baseline or iTraversal's result keys, for diffs across tools).
//
// Records, as written by the reference:
//   "Output n, cost Time: t ms" at the milestones {10^i < r} and r
//     (src/main.cu:1753-1761, include/for_generate.cuh:451-459);
//   "Running Time: t ms" and the two-column milestone table (src/main.cu:1709-1715);
//   JSON time record <dir>/<stem of the first .bin in the graph's directory><suffix>.json
//     = {"<k>": {"<r>": ms}}, pre-seeded with a 5-hour sentinel that stays when
//     the limit is not reached (include/timer.cuh:98-194); suffix ".gpu<k>".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gmk.h"

namespace fs = std::filesystem;

namespace {

// {"k": {"r": number}} with string keys; enough JSON for the time record.
using Record = std::map<std::string, std::map<std::string, double>>;

void skip_ws(const std::string& s, size_t& i) {
  while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t' || s[i] == ',')) ++i;
}
bool parse_str(const std::string& s, size_t& i, std::string& out) {
  skip_ws(s, i);
  if (i >= s.size() || s[i] != '"') return false;
  size_t j = s.find('"', i + 1);
  if (j == std::string::npos) return false;
  out = s.substr(i + 1, j - i - 1);
  i = j + 1;
  return true;
}
bool parse_record(const std::string& s, Record& rec) {
  size_t i = 0;
  skip_ws(s, i);
  if (i >= s.size() || s[i] != '{') return false;
  ++i;
  for (;;) {
    skip_ws(s, i);
    if (i < s.size() && s[i] == '}') return true;
    std::string k;
    if (!parse_str(s, i, k)) return false;
    skip_ws(s, i);
    if (i >= s.size() || s[i] != ':') return false;
    ++i;
    skip_ws(s, i);
    if (i >= s.size() || s[i] != '{') return false;
    ++i;
    for (;;) {
      skip_ws(s, i);
      if (i < s.size() && s[i] == '}') { ++i; break; }
      std::string r;
      if (!parse_str(s, i, r)) return false;
      skip_ws(s, i);
      if (i >= s.size() || s[i] != ':') return false;
      ++i;
      char* end = nullptr;
      const double v = strtod(s.c_str() + i, &end);
      if (end == s.c_str() + i) return false;
      i = end - s.c_str();
      rec[k][r] = v;
    }
  }
}
void write_record(const fs::path& p, const Record& rec) {
  std::ofstream o(p);
  o << "{\n";
  size_t a = 0;
  for (const auto& [k, m] : rec) {
    o << "    \"" << k << "\": {\n";
    size_t b = 0;
    for (const auto& [r, v] : m) {
      char buf[64];
      if (v == std::floor(v) && std::fabs(v) < 1e15) snprintf(buf, sizeof buf, "%.0f", v);
      else snprintf(buf, sizeof buf, "%.6f", v);
      o << "        \"" << r << "\": " << buf << (++b < m.size() ? ",\n" : "\n");
    }
    o << "    }" << (++a < rec.size() ? ",\n" : "\n");
  }
  o << "}";
}

// The reference names the record after the first .bin file of the graph's
// directory (include/timer.cuh:104-121).
fs::path record_path(const fs::path& graph, const std::string& suffix) {
  fs::path dir = graph.parent_path();
  if (dir.empty()) dir = ".";
  fs::path bin;
  for (const auto& e : fs::directory_iterator(dir))
    if (e.is_regular_file() && e.path().extension() == ".bin") { bin = fs::absolute(e.path()); break; }
  if (bin.empty()) bin = fs::absolute(graph);
  return bin.parent_path() / (bin.stem().string() + suffix + ".json");
}

// Reads the "tuning" object of an autotuner table: the arrays waves_per_cu and
// batch_states (7 numbers each), dedup_load_permille and check_form (4
// numbers; the built-in choice when absent).
bool read_tuning(const std::string& path, gmk_tuning& t) {
  std::ifstream in(path);
  if (!in) return false;
  std::stringstream ss;
  ss << in.rdbuf();
  const std::string s = ss.str();
  size_t at = s.find("\"tuning\"");
  if (at == std::string::npos) return false;
  auto numbers = [&](const char* key, uint32_t* out, int n) -> bool {
    size_t k = s.find(std::string("\"") + key + "\"", at);
    if (k == std::string::npos) return false;
    size_t i = s.find_first_of("[0123456789", k + strlen(key) + 2);
    if (i == std::string::npos) return false;
    if (s[i] == '[') ++i;
    for (int q = 0; q < n; ++q) {
      while (i < s.size() && (s[i] == ' ' || s[i] == ',' || s[i] == '\n')) ++i;
      char* end = nullptr;
      const unsigned long v = strtoul(s.c_str() + i, &end, 10);
      if (end == s.c_str() + i) return false;
      out[q] = (uint32_t)v;
      i = end - s.c_str();
    }
    return true;
  };
  t.version = GMK_TUNING_VERSION;
  const uint32_t form_default[4] = {0, 0, 0, 0};
  if (!numbers("check_form", t.check_form, 4)) memcpy(t.check_form, form_default, sizeof form_default);
  return numbers("waves_per_cu", t.waves_per_cu, 7) && numbers("batch_states", t.batch_states, 7) &&
         numbers("dedup_load_permille", &t.dedup_load_permille, 1);
}

// Result line formats of -q 1: the reference's own, and those of its two CPU
// enumerators, so outputs diff line by line across tools.
enum { FMT_GMKBPE = 0, FMT_BPPIVOT = 1, FMT_ITRAVERSAL = 2 };

void usage() {
  fprintf(stderr,
          "gmkbpe -f <graph.bin> [-k 1] [-r 10000000] [-d 0] [-q 0] [-m 0] [-t 1] "
          "[--mode plain|batched|learned] [--partition auto|global|seeded] [--tune table.json] "
          "[--format gmkbpe|bppivot|itraversal]\n");
}

}  // namespace

int main(int argc, char** argv) {
  std::string file;
  long long r = 10000000;
  int q = 0, m = 0, dev = 0;
  unsigned k = 1, theta = 1;
  int mode = GMK_MODE_BATCHED, part = GMK_PART_AUTO;
  std::string tune_path;
  int format = FMT_GMKBPE;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto val = [&]() -> const char* {
      if (i + 1 >= argc) { usage(); exit(1); }
      return argv[++i];
    };
    if (a == "-f" || a == "--file") file = val();
    else if (a == "-r" || a == "--r") r = atoll(val());
    else if (a == "-m") m = atoi(val());
    else if (a == "-q" || a == "--q") q = atoi(val());
    else if (a == "-d" || a == "--d") dev = atoi(val());
    else if (a == "-g" || a == "--g") (void)val();
    else if (a == "-k" || a == "--k") k = (unsigned)atoi(val());
    else if (a == "-t" || a == "--theta") theta = (unsigned)atoi(val());
    else if (a == "--mode") {
      std::string v = val();
      mode = v == "plain" ? GMK_MODE_PLAIN : v == "learned" ? GMK_MODE_LEARNED : GMK_MODE_BATCHED;
    } else if (a == "--partition") {
      std::string v = val();
      part = v == "global" ? GMK_PART_GLOBAL : v == "seeded" ? GMK_PART_SEEDED : GMK_PART_AUTO;
    } else if (a == "--tune") {
      tune_path = val();
    } else if (a == "--format") {
      std::string v = val();
      if (v == "gmkbpe") format = FMT_GMKBPE;
      else if (v == "bppivot") format = FMT_BPPIVOT;
      else if (v == "itraversal") format = FMT_ITRAVERSAL;
      else { fprintf(stderr, "unknown format %s\n", v.c_str()); usage(); return 1; }
    } else if (a == "-h" || a == "--help") { usage(); return 0; }
    else { fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return 1; }
  }
  if (file.empty()) { usage(); return 1; }
  // milestones {10^i < r} and r (src/main.cu:1753-1761)
  std::vector<unsigned long long> checkpoints;
  if (r > 0) {
    for (unsigned long long t = 1;; t *= 10) {
      if ((long long)t < r) checkpoints.push_back(t);
      else { checkpoints.push_back((unsigned long long)r); break; }
    }
  }
  gmk_ctx* ctx = nullptr;
  if (gmk_create(dev, &ctx) != GMK_OK) { fprintf(stderr, "Error: cannot open device %d\n", dev); return 1; }
  if (!tune_path.empty()) {
    gmk_tuning t{};
    if (!read_tuning(tune_path, t) || gmk_set_tuning(ctx, &t) != GMK_OK) {
      fprintf(stderr, "Error: cannot use tuning table %s: %s\n", tune_path.c_str(), gmk_last_error(ctx));
      gmk_destroy(ctx);
      return 1;
    }
  }
  {
    char name[256] = "";
    gmk_device_name(ctx, name, sizeof name);
    printf("Using GPU: %s, GPU id: %d\n", name, dev);  // src/main.cu:1768-1770
  }
  uint32_t n = 0, nl = 0;
  uint64_t nnz = 0;
  if (gmk_load_graph_bin(ctx, file.c_str(), &n, &nl, &nnz) != GMK_OK) {
    fprintf(stderr, "Error: %s\n", gmk_last_error(ctx));
    gmk_destroy(ctx);
    return 1;
  }
  printf("----------------------\n");
  printf("Graph size: %u, edge num: %llu, bipartite index: %u, k = %u, theta = %u\n", n, (unsigned long long)nnz, nl, k,
         theta);
  fflush(stdout);
  // time record pre-seeded with the 5-hour sentinel (include/timer.cuh:127-163)
  const fs::path rec_path = record_path(file, ".gpu" + std::to_string(k));
  Record rec;
  {
    std::ifstream in(rec_path);
    if (in) {
      std::stringstream ss;
      ss << in.rdbuf();
      if (!parse_record(ss.str(), rec)) rec.clear();
    }
  }
  const std::string kstr = std::to_string(k), rstr = std::to_string(r);
  rec[kstr][rstr] = 5.0 * 60 * 60 * 1000;
  write_record(rec_path, rec);

  gmk_params p{};
  p.k = k;
  p.theta = theta;
  p.limit = r > 0 ? (uint64_t)r : 0;
  p.mode = mode;
  p.partition = part;
  p.keep_results = q ? 1 : 0;
  p.verbose = m;
  gmk_stats st{};
  const int rc = gmk_enumerate(ctx, &p, &st);
  if (rc != GMK_OK) {
    fprintf(stderr, "Error: %s\n", gmk_last_error(ctx));
    gmk_destroy(ctx);
    return 1;
  }
  // milestone records from the progress log
  uint32_t np = 0;
  gmk_fetch_progress(ctx, nullptr, nullptr, 0, &np);
  std::vector<uint64_t> pc(np);
  std::vector<double> pm(np);
  gmk_fetch_progress(ctx, pc.data(), pm.data(), np, &np);
  std::vector<std::pair<unsigned long long, double>> times;
  for (unsigned long long cp : checkpoints) {
    for (uint32_t i = 0; i < np; ++i)
      if (pc[i] >= cp) { times.push_back({cp, pm[i]}); break; }
  }
  if (q) {
    uint64_t nw = 0;
    gmk_fetch_results(ctx, nullptr, 0, &nw);
    std::vector<uint32_t> w(nw);
    if (nw) gmk_fetch_results(ctx, w.data(), nw, &nw);
    uint64_t i = 0, c = 0;
    std::string line;
    while (i < nw) {
      const uint32_t a = w[i], b = w[i + 1 + a];
      const uint32_t* X = w.data() + i + 1;
      const uint32_t* Y = w.data() + i + 2 + a;
      ++c;
      if (format == FMT_BPPIVOT) {
        // "Biplex No.n:" then the sorted sides, right labels local to the right
        // side (baselines/MaximalBiPlex/src/biplex/biplex.cpp:812-813, 605-618)
        line = "Biplex No." + std::to_string(c) + ":\n  SL: ";
        for (uint32_t t = 0; t < a; ++t) line += std::to_string(X[t]) + ", ";
        line += "\n  SR: ";
        for (uint32_t t = 0; t < b; ++t) line += std::to_string(Y[t] - nl) + ", ";
      } else if (format == FMT_ITRAVERSAL) {
        // the result key "x&...y&..." of iTraversal's dedup tree: left then
        // right, each sorted, global IDs (SGMD22-MBPE/include/iTraversal.h:586-625)
        line.clear();
        for (uint32_t t = 0; t < a; ++t) line += std::to_string(X[t]) + "&";
        for (uint32_t t = 0; t < b; ++t) line += std::to_string(Y[t]) + "&";
      } else {
        // include/output.cuh:102-111, 136-144: X joined by spaces, each Y
        // followed by a space
        line = "Result " + std::to_string(c) + ": X: ";
        for (uint32_t t = 0; t < a; ++t) { line += std::to_string(X[t]); line += t + 1 < a ? " " : ""; }
        line += " | Y: ";
        for (uint32_t t = 0; t < b; ++t) { line += std::to_string(Y[t]); line += ' '; }
      }
      puts(line.c_str());
      i += 2 + a + b;
    }
  }
  for (const auto& [cp, ms] : times) printf("Output %llu, cost Time: %.6g ms\n", cp, ms);
  printf("Running Time: %.6g ms\n", st.ms);
  printf("\n===== \xe6\x80\xa7\xe8\x83\xbd\xe6\xb5\x8b\xe8\xaf\x95\xe6\x8a\xa5\xe5\x91\x8a =====\n");
  printf("\xe7\xbb\x93\xe6\x9e\x9c\xe6\x95\xb0\xe9\x87\x8f(r)\t\xe8\x80\x97\xe6\x97\xb6(ms)\n");
  for (const auto& [cp, ms] : times) printf("%llu\t\t%.6g\n", cp, ms);
  printf("========================\n");
  printf("Results: %llu%s\n", (unsigned long long)st.count, st.hit_limit ? " (limit reached)" : "");
  if (m) {
    static const char* names[4] = {"bb_expand", "thin_build_families", "thin_check", "partition_build"};
    for (int i = 0; i < 4; ++i)
      printf("  %-20s %10.3f ms  %8llu launches  %14llu algorithmic bytes\n", names[i], st.ms_kernel[i],
             (unsigned long long)st.launches_kernel[i], (unsigned long long)st.bytes_kernel[i]);
    printf("  states %llu  intersections %llu  partitions %llu\n", (unsigned long long)st.states,
           (unsigned long long)st.intersections, (unsigned long long)st.partitions);
  }
  if (r > 0 && st.count >= (uint64_t)r) {
    rec[kstr][rstr] = times.empty() ? st.ms : times.back().second;
    write_record(rec_path, rec);
  }
  gmk_destroy(ctx);
  return 0;
}
