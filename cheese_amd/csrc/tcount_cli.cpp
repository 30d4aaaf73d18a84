// `tcount SPEC [SPEC...] [--table K] [--max T] [--stats]`: the reference's T-count command line
// (src/tcount.cu:452-532) over the C ABI.  With one SPEC the stdout lines and exit code are the reference's: the
// table's level lines, the search's level lines, then the result line; exit 0 found, 1 none with T-count <= T,
// 2 bad spec, 3 unsupported, 4 overflow, 5 device.  With several SPECs the targets are searched together in one
// batch and every line of target i carries "target": i first; the exit code is the largest of the targets'.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cheese_hip.h"

static void level_line(const char *dir, const chz_tc_level &l, const char *prefix) {
  std::printf("{%s\"dir\": \"%s\", \"t\": %d, \"parents\": %ld, \"children\": %ld, \"cosets\": %ld, \"new\": %ld, "
              "\"seconds\": %.2f}\n",
              prefix, dir, l.t, l.parents, l.children, l.cosets, l.fresh, l.seconds);
}

int main(int argc, char **argv) {
  std::vector<std::string> paths;
  int K = 4, maxt = 16;
  bool stats = false;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--table") && i + 1 < argc) K = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--max") && i + 1 < argc) maxt = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--stats")) stats = true;
    else paths.push_back(argv[i]);
  }
  if (paths.empty()) {
    std::fprintf(stderr, "usage: tcount SPEC [SPEC...] [--table K] [--max T] [--stats]\n"
                         "optimal T-count up to global phase; exit 0 found, 1 none with T-count <= T\n");
    return 3;
  }
  std::vector<std::string> texts;
  for (const std::string &p : paths) {
    std::ifstream f(p);
    if (!f) { std::fprintf(stderr, "cannot read %s\n", p.c_str()); return 2; }
    std::stringstream ss;
    ss << f.rdbuf();
    texts.push_back(ss.str());
  }
  chz_tc *t = nullptr;
  int rc = chz_tc_table(texts[0].data(), texts[0].size(), K, 0, &t);
  if (rc) { std::fprintf(stderr, "%s\n", chz_tc_last_error()); return rc; }
  std::vector<chz_tc_level> tl(64);
  const int nl = chz_tc_table_levels(t, tl.data(), 64);
  for (int l = 0; l < nl; l++) level_line("table", tl[l], "");
  std::fprintf(stderr, "table: %lld cosets with T-count <= %d\n", chz_tc_table_size(t), chz_tc_table_depth(t));
  const int n = (int)texts.size();
  std::vector<const char *> sp(n);
  std::vector<size_t> ln(n);
  for (int i = 0; i < n; i++) { sp[i] = texts[i].data(); ln[i] = texts[i].size(); }
  std::vector<chz_tc_result> res(n);
  rc = chz_tc_search(t, sp.data(), ln.data(), n, maxt, res.data());
  if (rc) { std::fprintf(stderr, "%s\n", chz_tc_last_error()); chz_tc_free(t); return rc; }
  int worst = 0;
  for (int i = 0; i < n; i++) {
    const chz_tc_result &r = res[i];
    char pre[32] = "";
    if (n > 1) std::snprintf(pre, sizeof pre, "\"target\": %d, ", i);
    for (int l = 0; l < r.nlevels; l++) level_line("search", r.levels[l], pre);
    if (r.rc == 0) {
      std::printf("{%s\"tcount\": %d, \"rotations\": [", pre, r.tcount);
      for (int k = 0; k < r.tcount; k++) std::printf("%s[%d, %d]", k ? ", " : "", r.rotations[2 * k], r.rotations[2 * k + 1]);
      std::printf("], \"split\": [%d, %d], \"seconds\": %.2f}\n", r.split[0], r.split[1], r.seconds);
    } else if (r.rc == 1) {
      std::printf("{%s\"tcount_above\": %d, \"seconds\": %.2f}\n", pre, maxt, r.seconds);
    } else {
      std::fprintf(stderr, "target %d: exit %d\n", i, r.rc);
    }
    if (r.rc > worst) worst = r.rc;
  }
  if (stats) {
    chz_kstat ks[16];
    const int nk = chz_tc_kernel_stats(t, ks, 16);
    for (int k = 0; k < nk; k++)
      std::fprintf(stderr, "  %-12s %6ld launches %10.3f ms\n", ks[k].name, ks[k].launches, ks[k].ms);
    double d[4];
    chz_tc_dedupe_counts(t, d);
    std::fprintf(stderr, "  dedupe: %.0f runs, %.0f children visited, %.0f exact comparisons, %.0f duplicates\n",
                 d[0], d[1], d[2], d[3]);
  }
  chz_tc_free(t);
  return worst;
}
