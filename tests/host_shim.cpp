// Test-only C entry points over the product's host spec code (cheese_amd/csrc/spec.cpp), built by
// tests/test_host_spec.py with g++ so that the host logic is checked without a GPU.
#include <string>

#include "../cheese_amd/csrc/spec.h"

extern "C" int shim_gate_info(const char *spec, long len, int g, int *out, int cap) {
  chz::HostSpec s;
  std::string err;
  if (chz::parse_spec(std::string(spec, (size_t)len), s, err) || g < 0 || g >= s.G()) return -1;
  const int nc = (int)s.cutmask.size();
  if (cap < 2 + nc) return -1;
  out[0] = chz::gate_is_clifford(s.gates[g]) ? 1 : 0;
  out[1] = chz::gate_ptm_sde(s.gates[g]);
  for (int i = 0; i < nc; i++) out[2 + i] = chz::gate_cut_rank(s.gates[g], s.cutmask[i]);
  return nc;
}
