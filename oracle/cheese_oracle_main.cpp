// ORACLE command line (test infrastructure only): `cheese-oracle SPEC [--depth D] [--no-dedupe] [--no-sym]`
// prints the same stdout lines as the reference binary (src/cheese.cu:1851-1895) and exits with its code.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

extern "C" int orc_run(const char *spec, long len, int depth, int dedupe, int sym, int threads, char *out, long cap,
                       long *stats);

int main(int argc, char **argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: cheese-oracle SPEC [--depth D] [--no-dedupe] [--no-sym] [--stats]\n"); return 3; }
  std::ifstream f(argv[1]);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  int depth = -1, dedupe = 1, sym = 1, stats = 0;
  for (int i = 2; i < argc; i++) {
    if (!std::strcmp(argv[i], "--depth") && i + 1 < argc) depth = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--no-dedupe")) dedupe = 0;
    else if (!std::strcmp(argv[i], "--no-sym")) sym = 0;
    else if (!std::strcmp(argv[i], "--stats")) stats = 1;
  }
  std::vector<char> out(1 << 20);
  long st[3] = {0, 0, 0};
  const int rc = orc_run(text.data(), (long)text.size(), depth, dedupe, sym, 0, out.data(), (long)out.size(), st);
  std::fputs(out.data(), stdout);
  if (stats) std::fprintf(stderr, "cheap %ld tbound %ld hashed %ld\n", st[0], st[1], st[2]);
  return rc;
}
