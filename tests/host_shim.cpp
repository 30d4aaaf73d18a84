// Test-only C entry points over the product's host spec code (cheese_amd/csrc/spec.cpp), built by
// tests/test_host_spec.py with g++ so that the host logic is checked without a GPU.
#include <string>
#include <vector>

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

// goal pre-image of candidate g: its sqrt2 exponent (or -1) and N x N entries as (a, b, c, d) in out
extern "C" int shim_goal_preimage(const char *spec, long len, int g, long long *out, int cap) {
  chz::HostSpec s;
  std::string err;
  if (chz::parse_spec(std::string(spec, (size_t)len), s, err) || g < 0 || g >= s.G()) return -2;
  chz::HostMat X;
  if (!chz::host_goal_preimage(s, g, X)) return -1;
  if (cap < 4 * (int)X.e.size()) return -2;
  for (size_t i = 0; i < X.e.size(); i++) {
    out[4 * i] = X.e[i].a; out[4 * i + 1] = X.e[i].b; out[4 * i + 2] = X.e[i].c; out[4 * i + 3] = X.e[i].d;
  }
  return X.k;
}

// whether candidates g1 and g2 of the spec give children of equal T bound from every residual (same_t_bound)
extern "C" int shim_same_t_bound(const char *spec, long len, int g1, int g2) {
  chz::HostSpec s;
  std::string err;
  if (chz::parse_spec(std::string(spec, (size_t)len), s, err) || g1 < 0 || g2 < 0 || g1 >= s.G() || g2 >= s.G())
    return -1;
  return chz::same_t_bound(s.gates[g1], s.gates[g2]) ? 1 : 0;
}

// whether every candidate has a goal pre-image (the condition of the found-level skip), after multiplying the
// goal's entries by `scale` (a large scale makes pre-images leave int32, so some candidates have none)
extern "C" int shim_goal_skip(const char *spec, long len, long long scale) {
  chz::HostSpec s;
  std::string err;
  if (chz::parse_spec(std::string(spec, (size_t)len), s, err)) return -2;
  for (chz::Zq &z : s.goal) { z.a *= scale; z.b *= scale; z.c *= scale; z.d *= scale; }
  std::vector<chz::HostMat> X;
  std::vector<char> have;
  return chz::goal_preimages(s, X, have) ? 1 : 0;
}
