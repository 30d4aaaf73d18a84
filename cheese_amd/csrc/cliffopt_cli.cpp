// `cliffopt JOB [--beam B]`: the reference's Clifford-segment optimiser (src/clifford.cu:285-385) over the C ABI:
// the Clifford table of the job's qubit count, the T-placement beam search, the same stdout.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cheese_hip.h"

int main(int argc, char **argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: cliffopt JOB [--beam B]\n"); return 3; }
  std::ifstream f(argv[1]);
  if (!f) { std::fprintf(stderr, "cannot read %s\n", argv[1]); return 2; }
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string job = ss.str();
  int beam = 16;
  for (int i = 2; i < argc; i++)
    if (!std::strcmp(argv[i], "--beam") && i + 1 < argc) beam = std::max(1, std::min(255, std::atoi(argv[++i])));
  int n = 0;
  if (std::sscanf(job.c_str(), "%d", &n) != 1 || n < 1 || n > 3) { std::fprintf(stderr, "bad job file\n"); return 2; }
  chz_cliff *c = nullptr;
  int rc = chz_cliff_build(n, 0, &c);
  if (rc) { std::fprintf(stderr, "%s\n", chz_cliff_last_error()); return rc; }
  std::vector<char> out(1 << 22);
  rc = chz_cliff_opt(c, job.data(), job.size(), beam, out.data(), out.size());
  if (rc) std::fprintf(stderr, "%s\n", chz_cliff_last_error());
  else std::fputs(out.data(), stdout);
  chz_cliff_free(c);
  return rc;
}
