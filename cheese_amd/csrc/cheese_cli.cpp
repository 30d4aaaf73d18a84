// `cheese SPEC [--depth D] [--no-dedupe] [--chunk PARENTS] [--mitm] [--no-sym] [--table K] [--profile]`: the reference binary's
// command line (src/cheese.cu:1851-1895) over the C ABI: same stdout lines, same exit codes.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/cheese_hip.h"

static void on_level(const chz_level *l, void *) {
  if (l->dir == 4) {       // the meet-in-the-middle join (src/cheese.cu:1714-1715)
    std::printf("{\"dir\": \"join\", \"forward\": %ld, \"backward\": %ld, \"hash_matches\": %ld, \"seconds\": %.2f}\n",
                l->parents, l->prefixes, l->distinct, l->seconds);
    std::fflush(stdout);
    return;
  }
  static const char *dirs[] = {"", "\"dir\": \"table\", ", "\"dir\": \"fwd\", ", "\"dir\": \"bwd\", "};
  std::printf("{%s\"t\": %d, \"parents\": %ld, \"prefixes\": %ld, \"cut_w3\": %ld, \"cut_lb\": %ld, \"survivors\": %ld, "
              "\"distinct\": %ld, \"proved_unsat\": %s, \"seconds\": %.2f}\n",
              l->dir >= 1 && l->dir <= 3 ? dirs[l->dir] : "", l->t, l->parents, l->prefixes, l->cut_w3, l->cut_lb,
              l->survivors, l->distinct, l->proved_unsat ? "true" : "false", l->seconds);
  std::fflush(stdout);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: cheese SPEC [--depth D] [--no-dedupe] [--chunk PARENTS] [--mitm] [--no-sym] [--table K] [--profile]\n"
                         "exit 0: circuit found, 1: no circuit of <= depth gates (proof)\n");
    return 3;
  }
  std::ifstream f(argv[1]);
  if (!f) { std::fprintf(stderr, "cannot read %s\n", argv[1]); return 2; }
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  chz_opts o{-1, 1, 1, 0, 0, 0, 0, 0};
  for (int i = 2; i < argc; i++) {
    if (!std::strcmp(argv[i], "--depth") && i + 1 < argc) o.depth = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--no-dedupe")) o.dedupe = 0;
    else if (!std::strcmp(argv[i], "--chunk") && i + 1 < argc) o.chunk = std::atol(argv[++i]);
    else if (!std::strcmp(argv[i], "--profile")) o.profile = 1;
    else if (!std::strcmp(argv[i], "--no-sym")) o.sym = 0;
    else if (!std::strcmp(argv[i], "--mitm")) o.mitm = 1;
    else if (!std::strcmp(argv[i], "--table") && i + 1 < argc) o.table = std::atoi(argv[++i]);
  }
  chz_ctx *ctx = nullptr;
  int rc = chz_open(text.data(), text.size(), &o, &ctx);
  if (rc) { std::fprintf(stderr, "%s\n", chz_last_error()); return rc; }
  std::fprintf(stderr, "cheese: n %d, %d candidates, depth %d, dedupe %d\n", chz_num_qubits(ctx), chz_num_gates(ctx),
               chz_depth(ctx), o.dedupe);
  std::fprintf(stderr, "symmetry classes: %d group elements\n", chz_sym_size(ctx));
  chz_result r{};
  rc = chz_run(ctx, on_level, nullptr, &r);
  if (rc == 0) {
    if (r.len == 0) std::printf("{\"found\": 0, \"circuit\": []}\n");
    else {
      // a found count of 0 prints as 1, as the reference's print_circuit does (src/cheese.cu:1659)
      std::printf("{\"t\": %d, \"found\": %ld, \"circuit\": [", r.len, r.found ? r.found : 1L);
      for (int i = 0; i < r.len; i++) std::printf("%s\"%s\"", i ? ", " : "", chz_gate_name(ctx, r.circuit[i]));
      if (r.split_kind) std::printf("], \"%s\": [%d, %d", r.split_kind == 1 ? "table" : "mitm", r.split[0], r.split[1]);
      std::printf("], \"seconds\": %.2f}\n", r.seconds);
      if (chz_verify(ctx, r.circuit, r.len) != 1) {
        std::fprintf(stderr, "internal error: the circuit does not reproduce the target\n");
        rc = 5;
      }
    }
  } else if (rc == 1 && r.split_kind == 1) {        // the table mode's proof line (src/cheese.cu:1840)
    std::printf("{\"depth\": %d, \"proved_unsat\": true, \"table\": %d}\n", chz_depth(ctx), r.split[1]);
  } else if (rc == 1 && r.split_kind == 2) {        // the join's proof line (src/cheese.cu:1721, :1758)
    std::printf("{\"depth\": %d, \"proved_unsat\": true, \"mitm\": [%d, %d]}\n", chz_depth(ctx), r.split[0], r.split[1]);
  } else if (rc >= 2) {
    std::fprintf(stderr, "%s\n", chz_last_error());
  }
  if (o.profile) {
    const char *nm[16];
    double s[16];
    const int np = chz_profile(ctx, nm, s, 16);
    for (int i = 0; i < np; i++) std::fprintf(stderr, "  %-9s %.3f s\n", nm[i], s[i]);
  }
  chz_close(ctx);
  return rc;
}
