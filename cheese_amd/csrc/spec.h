// Host-side spec model of the gate-count search: the parsed .spec, the symmetry group and exact host
// arithmetic for checks that run once per search (depth-0 test, circuit verification).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace chz {

constexpr int kDmax = 48, kGmax = 64, kLmax = 8, kCutMax = 16;
constexpr uint32_t kPrime = 998244353u;   // p = 1 mod 8: w has an image in F_p (src/cheese.cu:57)

// a + b w + c w^2 + d w^3 with w = exp(i pi / 4)
struct Zq { int64_t a, b, c, d; };

struct HostGate {
  std::string name;       // L_<gate>_<q>[_<q>...]
  int nq = 0, q[3] = {0, 0, 0};
  int gk = 0, ik = 0;     // sqrt2 exponents of the forward and inverse local matrices
  std::vector<Zq> fwd, inv;   // L x L, L = 2^nq
};

struct HostSpec {
  int n = 0, depth = 0, standard = 0, lite = 0, tdiv = 1, states = 0;
  int tk = 0, gk = 0;
  bool truncated = false;        // the token stream ended inside a field
  std::vector<HostGate> gates;
  std::vector<uint8_t> bad2;     // G x G
  std::vector<uint32_t> bad3;    // bitset over G^3
  std::vector<Zq> target, goal;  // N x N
  std::vector<int> cutmask, cutsize, rad, cxtable;
  int G() const { return (int)gates.size(); }
  int N() const { return 1 << n; }
};

// Parses .spec text (format: src/cheese.cu:1149-1226, written by python/cheese/spec.py:356-415).
// Returns 0, or 2 with a message in err for a malformed spec.
int parse_spec(const std::string &text, HostSpec &out, std::string &err);

// Symmetry group: qubit relabellings x optional conjugation that map every candidate to a candidate
// (src/cheese.cu:1383-1429).  bperm: nsym x N basis maps, conj: nsym flags, gmap: candidate maps.
struct SymGroup {
  std::vector<uint8_t> bperm, conj;
  std::vector<std::vector<int>> gmap;
  int size() const { return (int)conj.size(); }
};
SymGroup build_symmetry(const HostSpec &s, bool on);

// exact host matrices (entries / sqrt2^k), used for the depth-0 test and circuit verification
struct HostMat { std::vector<Zq> e; int k = 0; bool ovf = false; };
void host_apply_inverse(HostMat &S, const HostGate &g, int n);   // S <- S g^-1, normalised
bool host_equal(const HostMat &S, const std::vector<Zq> &M, int k);
// X with X g^-1 = goal, in normal form: the one residual from which candidate gi reaches the goal.  False when
// the product check fails or a coefficient leaves int32.
bool host_goal_preimage(const HostSpec &s, int gi, HostMat &X);
// the pre-images of every candidate (X[g], with have[g] = 0 where host_goal_preimage fails); returns true when
// every candidate has one, which the search's found-level skip requires (a parent without a goal child is passed
// over only if no candidate can reach the goal from it, which is decided by comparing with every pre-image)
bool goal_preimages(const HostSpec &s, std::vector<HostMat> &X, std::vector<char> &have);

// two specs share everything the search kernels read except the target and the depth: qubit count, mode,
// candidates and their matrices, windows, cuts, cx table, bound flags and goal (searches of such specs can run
// as one batch)
bool same_gate_set(const HostSpec &a, const HostSpec &b);

// exact: the gate maps every Pauli operator to a multiple of a Pauli operator under conjugation
bool gate_is_clifford(const HostGate &g);
// largest sqrt2-denominator exponent of the gate's Pauli transfer matrix (0: Clifford; 1: T), -1 if malformed
int gate_ptm_sde(const HostGate &g);
// operator-Schmidt rank over F_p of the gate across cut mask A (1 when all its qubits are on one side)
int gate_cut_rank(const HostGate &g, int A);
// true when candidates a and b act on the same qubits and a b^-1 is a Clifford: then for every residual S the
// children S a^-1 and S b^-1 = (S a^-1)(a b^-1) have equal T bounds (a Clifford's transfer matrix is a signed
// permutation).  T and T^dagger on one qubit are such a pair: T T = S.
bool same_t_bound(const HostGate &a, const HostGate &b);

}  // namespace chz
