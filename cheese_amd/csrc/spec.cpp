// .spec reader, symmetry group and exact host arithmetic for the gate-count search.
#include "spec.h"

#include <algorithm>
#include <cstdlib>
#include <sstream>

namespace chz {

static Zq zq_mul(Zq x, Zq y) {
  return {x.a * y.a - x.b * y.d - x.c * y.c - x.d * y.b, x.a * y.b + x.b * y.a - x.c * y.d - x.d * y.c,
          x.a * y.c + x.b * y.b + x.c * y.a - x.d * y.d, x.a * y.d + x.b * y.c + x.c * y.b + x.d * y.a};
}
static bool zq_zero(const Zq &x) { return !(x.a | x.b | x.c | x.d); }
static bool zq_eq(const Zq &x, const Zq &y) { return x.a == y.a && x.b == y.b && x.c == y.c && x.d == y.d; }
static bool zq_sqrt2_divides(const Zq &x) { return !((x.a - x.c) & 1) && !((x.b - x.d) & 1); }
static Zq zq_div_sqrt2(const Zq &x) { return {(x.b - x.d) / 2, (x.a + x.c) / 2, (x.b + x.d) / 2, (x.c - x.a) / 2}; }

int parse_spec(const std::string &text, HostSpec &s, std::string &err) {
  std::istringstream in(text);
  std::string tok;
  if (!(in >> tok) || tok != "cheese-spec") { err = "not a spec"; return 2; }
  int version = 0;
  in >> version;
  int G = 0;
  auto read_z = [&](Zq &z) { in >> z.a >> z.b >> z.c >> z.d; };
  while (in >> tok) {
    if (tok == "n") in >> s.n;
    else if (tok == "mode" || tok == "phase") {
      std::string v;
      in >> v;
      if (v == "states") s.states = 1;
      else if (v != "unitary" && v != "exact") { err = "unsupported " + tok + " " + v; return 2; }
    } else if (tok == "pairs") { int m; in >> m; }
    else if (tok == "start") {
      if (s.n < 1 || s.n > 6) { err = "n = " + std::to_string(s.n) + " out of range"; return 3; }
      int k; in >> k;
      const int N = 1 << s.n;
      for (int i = 0; i < N * N; i++) { Zq z; read_z(z); }
    } else if (tok == "goal") {
      int m; in >> m;
      const int w = 2 + (int)s.cutmask.size();
      for (int i = 0; i < m * w; i++) { int v; in >> v; }
    } else if (tok == "depth") in >> s.depth;
    else if (tok == "standard") in >> s.standard;
    else if (tok == "lite") in >> s.lite;
    else if (tok == "tdiv") in >> s.tdiv;
    else if (tok == "gates") {
      in >> G;
      if (G > kGmax) { err = "too many gates"; return 2; }
      if (G < 0) { err = "bad gate count"; return 2; }
      s.gates.assign(G, HostGate{});
      for (HostGate &g : s.gates) {
        in >> g.name >> g.nq;
        if (g.nq < 1 || g.nq > 3) { err = "bad gate arity"; return 2; }
        for (int j = 0; j < g.nq; j++) in >> g.q[j];
        const int L = 1 << g.nq;
        in >> g.gk;
        g.fwd.resize(L * L);
        for (Zq &z : g.fwd) read_z(z);
        in >> g.ik;
        g.inv.resize(L * L);
        for (Zq &z : g.inv) read_z(z);
      }
    } else if (tok == "bad2") {
      int m; in >> m;
      s.bad2.assign((size_t)G * G, 0);
      for (int i = 0; i < m; i++) {
        int a, b; in >> a >> b;
        if (a < 0 || b < 0 || a >= G || b >= G) { err = "bad2 index"; return 2; }
        s.bad2[(size_t)a * G + b] = 1;
      }
    } else if (tok == "bad3") {
      int m; in >> m;
      s.bad3.assign(((size_t)G * G * G + 31) / 32, 0);
      for (int i = 0; i < m; i++) {
        int a, b, c; in >> a >> b >> c;
        if (a < 0 || b < 0 || c < 0 || a >= G || b >= G || c >= G) { err = "bad3 index"; return 2; }
        const uint32_t t = ((uint32_t)a * G + b) * G + c;
        s.bad3[t >> 5] |= 1u << (t & 31);
      }
    } else if (tok == "target") {
      if (s.n < 1 || s.n > 6) { err = "n = " + std::to_string(s.n) + " out of range"; return 3; }
      int k; in >> k;
      const int N = 1 << s.n;
      std::vector<Zq> M(N * N);
      for (Zq &z : M) read_z(z);
      if (s.states) { s.goal = M; s.gk = k; }
      else {
        s.target = M; s.tk = k; s.gk = 0;
        s.goal.assign(N * N, Zq{0, 0, 0, 0});
        for (int i = 0; i < N; i++) s.goal[i * N + i] = Zq{1, 0, 0, 0};
      }
    } else if (tok == "cuts") {
      int m; in >> m;
      for (int i = 0; i < m; i++) {
        int sz, mask = 0; in >> sz;
        for (int j = 0; j < sz; j++) { int q; in >> q; mask |= 1 << q; }
        s.cutmask.push_back(mask); s.cutsize.push_back(sz); s.rad.push_back(2 * sz + 1);
      }
      if (m > kCutMax) { err = "too many cuts"; return 2; }
    } else if (tok == "cxtable") {
      int m; in >> m;
      s.cxtable.resize(m);
      for (int &v : s.cxtable) in >> v;
    } else if (tok == "end") break;
    else { err = "unknown spec field " + tok; return 2; }
  }
  s.truncated = in.fail() && !in.eof();
  if (s.bad2.empty()) s.bad2.assign((size_t)G * G, 0);
  if (s.bad3.empty()) s.bad3.assign(((size_t)G * G * G + 31) / 32, 0);
  if (s.tdiv < 1) s.tdiv = 1;
  return 0;
}

// the candidate that a relabelling perm (and conjugation cj) maps candidate g to, or -1 (src/cheese.cu:1383-1402)
static int image_of(const HostSpec &s, const std::vector<int> &perm, bool cj, int g) {
  const std::string &nm = s.gates[g].name;
  std::vector<std::string> part;
  size_t a = 0, b;
  while ((b = nm.find('_', a)) != std::string::npos) { part.push_back(nm.substr(a, b - a)); a = b + 1; }
  part.push_back(nm.substr(a));
  if (part.size() < 2) return -1;
  std::string kind = part[1];
  if (cj) {
    static const char *swap[4][2] = {{"t", "tdg"}, {"tdg", "t"}, {"s", "sdg"}, {"sdg", "s"}};
    bool swapped = false;
    for (auto &p : swap) if (kind == p[0]) { kind = p[1]; swapped = true; break; }
    if (!swapped && kind != "x" && kind != "z" && kind != "h" && kind != "cx" && kind != "cz" && kind != "ccx") return -1;
  }
  std::vector<int> qs;
  for (size_t i = 2; i < part.size(); i++) {
    const int q = std::atoi(part[i].c_str());
    if (q < 0 || q >= (int)perm.size()) return -1;
    qs.push_back(perm[q]);
  }
  if (kind == "ccx" && qs.size() == 3 && qs[0] > qs[1]) std::swap(qs[0], qs[1]);   // controls unordered
  std::string want = "L_" + kind;
  for (int q : qs) want += "_" + std::to_string(q);
  for (int h = 0; h < s.G(); h++) if (s.gates[h].name == want) return h;
  return -1;
}

SymGroup build_symmetry(const HostSpec &s, bool on) {
  SymGroup grp;
  const int n = s.n, N = 1 << n;
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  std::vector<std::vector<int>> perms;
  do perms.push_back(perm); while (on && std::next_permutation(perm.begin(), perm.end()));
  for (int cj = 0; cj < (on ? 2 : 1); cj++)
    for (const auto &pm : perms) {
      std::vector<int> gm(s.G());
      bool closed = true;
      for (int g = 0; g < s.G() && closed; g++) closed = (gm[g] = image_of(s, pm, cj, g)) >= 0;
      if (!closed) continue;
      grp.conj.push_back((uint8_t)cj);
      grp.gmap.push_back(gm);
      for (int x = 0; x < N; x++) {
        int y = 0;
        for (int i = 0; i < n; i++) if ((x >> i) & 1) y |= 1 << pm[i];
        grp.bperm.push_back((uint8_t)y);
      }
    }
  return grp;
}

void host_apply_inverse(HostMat &S, const HostGate &g, int n) {
  const int N = 1 << n, L = 1 << g.nq;
  int mask = 0;
  for (int j = 0; j < g.nq; j++) mask |= 1 << g.q[j];
  std::vector<Zq> out(S.e.size());
  for (int r = 0; r < N; r++)
    for (int c = 0; c < N; c++) {
      int lc = 0;
      for (int j = 0; j < g.nq; j++) lc |= ((c >> g.q[j]) & 1) << j;
      const int base = c & ~mask;
      Zq acc{0, 0, 0, 0};
      for (int m = 0; m < L; m++) {
        const Zq gm = g.inv[m * L + lc];
        if (zq_zero(gm)) continue;
        int x = base;
        for (int j = 0; j < g.nq; j++) if ((m >> j) & 1) x |= 1 << g.q[j];
        const Zq p = zq_mul(S.e[r * N + x], gm);
        acc = {acc.a + p.a, acc.b + p.b, acc.c + p.c, acc.d + p.d};
      }
      if (acc.a != (int32_t)acc.a || acc.b != (int32_t)acc.b || acc.c != (int32_t)acc.c || acc.d != (int32_t)acc.d) S.ovf = true;
      out[r * N + c] = acc;
    }
  S.e.swap(out);
  S.k += g.ik;
  while (S.k > 0 && std::all_of(S.e.begin(), S.e.end(), zq_sqrt2_divides)) {
    for (Zq &z : S.e) z = zq_div_sqrt2(z);
    S.k--;
  }
}

bool host_equal(const HostMat &S, const std::vector<Zq> &M, int k) {
  if (S.k != k || S.e.size() != M.size()) return false;
  for (size_t i = 0; i < M.size(); i++) if (!zq_eq(S.e[i], M[i])) return false;
  return true;
}

bool host_goal_preimage(const HostSpec &s, int gi, HostMat &X) {
  const HostGate &g = s.gates[gi];
  const int N = s.N(), L = 1 << g.nq;
  if ((int)s.goal.size() != N * N || (int)g.fwd.size() != L * L) return false;
  int mask = 0;
  for (int j = 0; j < g.nq; j++) mask |= 1 << g.q[j];
  // X = goal G: entry (r, c) = sum_m goal[r][x_m] fwd[m][lc], x_m = c with the gate's bits set to m
  X.e.assign((size_t)N * N, Zq{0, 0, 0, 0});
  X.k = s.gk + g.gk;
  X.ovf = false;
  for (int r = 0; r < N; r++)
    for (int c = 0; c < N; c++) {
      int lc = 0;
      for (int j = 0; j < g.nq; j++) lc |= ((c >> g.q[j]) & 1) << j;
      Zq acc{0, 0, 0, 0};
      for (int m = 0; m < L; m++) {
        int x = c & ~mask;
        for (int j = 0; j < g.nq; j++) if ((m >> j) & 1) x |= 1 << g.q[j];
        const Zq p = zq_mul(s.goal[r * N + x], g.fwd[m * L + lc]);
        acc = {acc.a + p.a, acc.b + p.b, acc.c + p.c, acc.d + p.d};
      }
      X.e[r * N + c] = acc;
    }
  while (X.k > 0 && std::all_of(X.e.begin(), X.e.end(), zq_sqrt2_divides)) {
    for (Zq &z : X.e) z = zq_div_sqrt2(z);
    X.k--;
  }
  for (const Zq &z : X.e)
    if (z.a != (int32_t)z.a || z.b != (int32_t)z.b || z.c != (int32_t)z.c || z.d != (int32_t)z.d) return false;
  HostMat Y = X;
  host_apply_inverse(Y, g, s.n);
  return !Y.ovf && host_equal(Y, s.goal, s.gk);
}

bool goal_preimages(const HostSpec &s, std::vector<HostMat> &X, std::vector<char> &have) {
  X.assign(s.G(), HostMat{});
  have.assign(s.G(), 0);
  bool all = true;
  for (int g = 0; g < s.G(); g++) {
    have[g] = s.N() * s.N() <= 64 && host_goal_preimage(s, g, X[g]);
    all = all && have[g];
  }
  return all;
}

static bool zq_vec_eq(const std::vector<Zq> &a, const std::vector<Zq> &b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); i++)
    if (a[i].a != b[i].a || a[i].b != b[i].b || a[i].c != b[i].c || a[i].d != b[i].d) return false;
  return true;
}

bool same_gate_set(const HostSpec &a, const HostSpec &b) {
  if (a.n != b.n || a.states != b.states || a.standard != b.standard || a.lite != b.lite || a.tdiv != b.tdiv ||
      a.gk != b.gk || a.G() != b.G() || a.bad2 != b.bad2 || a.bad3 != b.bad3 || a.cutmask != b.cutmask ||
      a.cutsize != b.cutsize || a.rad != b.rad || a.cxtable != b.cxtable || !zq_vec_eq(a.goal, b.goal))
    return false;
  for (int g = 0; g < a.G(); g++) {
    const HostGate &x = a.gates[g], &y = b.gates[g];
    if (x.nq != y.nq || x.gk != y.gk || x.ik != y.ik || !zq_vec_eq(x.fwd, y.fwd) || !zq_vec_eq(x.inv, y.inv)) return false;
    for (int j = 0; j < x.nq; j++) if (x.q[j] != y.q[j]) return false;
  }
  return true;
}

// Pauli transfer matrix of a gate, R_ab = Tr(P_a g P_b g^dagger) / L with g = M / sqrt2^gk: each entry is
// t_ab / sqrt2^(2 (nq + gk)) with t_ab = Tr(P_a M P_b M^dagger) in Z[w].  Returns the largest sqrt2-denominator
// exponent over the entries (0 exactly for Clifford gates: an orthogonal matrix over Z[sqrt2] whose Galois
// conjugate is orthogonal too has entries 0, +-1 only), or -1 when M is not L x L.
static int sqrt2_valuation(Zq x) {
  int v = 0;
  while (zq_sqrt2_divides(x)) { x = zq_div_sqrt2(x); v++; }
  return v;
}
static Zq zq_conj(const Zq &x) { return {x.a, -x.d, -x.c, -x.b}; }
int gate_ptm_sde(const HostGate &g) {
  const int L = 1 << g.nq;
  const std::vector<Zq> &M = g.fwd;
  if ((int)M.size() != L * L) return -1;
  const int scale = 2 * (g.nq + g.gk);
  std::vector<Zq> MB(L * L), C(L * L);
  int sde = 0;
  for (int bx = 0; bx < L; bx++)
    for (int bz = 0; bz < L; bz++) {
      // (M B)[r][c] = M[r][c ^ bx] (-1)^(bz . c): B |c> = (-1)^(bz . c) |c ^ bx>
      for (int r = 0; r < L; r++)
        for (int c = 0; c < L; c++) {
          Zq v = M[r * L + (c ^ bx)];
          if (__builtin_popcount(bz & c) & 1) v = {-v.a, -v.b, -v.c, -v.d};
          MB[r * L + c] = v;
        }
      for (int r = 0; r < L; r++)       // C = (M B) M^dagger
        for (int c = 0; c < L; c++) {
          Zq acc{0, 0, 0, 0};
          for (int j = 0; j < L; j++) {
            const Zq t = zq_mul(MB[r * L + j], zq_conj(M[c * L + j]));
            acc = {acc.a + t.a, acc.b + t.b, acc.c + t.c, acc.d + t.d};
          }
          C[r * L + c] = acc;
        }
      for (int ax = 0; ax < L; ax++)
        for (int az = 0; az < L; az++) {
          Zq tr{0, 0, 0, 0};          // Tr(A^dagger C) = sum_c <c ^ ax| C |c> (-1)^(az . c), up to a unit
          for (int c = 0; c < L; c++) {
            Zq v = C[(c ^ ax) * L + c];
            if (__builtin_popcount(az & c) & 1) v = {-v.a, -v.b, -v.c, -v.d};
            tr = {tr.a + v.a, tr.b + v.b, tr.c + v.c, tr.d + v.d};
          }
          if (zq_zero(tr)) continue;
          const int need = scale - sqrt2_valuation(tr);
          if (need > sde) sde = need;
        }
    }
  return sde;
}

bool gate_is_clifford(const HostGate &g) { return gate_ptm_sde(g) == 0; }

bool same_t_bound(const HostGate &a, const HostGate &b) {
  if (a.nq != b.nq) return false;
  for (int j = 0; j < a.nq; j++) if (a.q[j] != b.q[j]) return false;
  const int L = 1 << a.nq;
  if ((int)a.fwd.size() != L * L || (int)b.inv.size() != L * L) return false;
  HostGate c;
  c.nq = a.nq;
  c.gk = a.gk + b.ik;
  c.fwd.assign(L * L, Zq{0, 0, 0, 0});
  for (int r = 0; r < L; r++)
    for (int col = 0; col < L; col++) {
      Zq acc{0, 0, 0, 0};
      for (int j = 0; j < L; j++) {
        const Zq t = zq_mul(a.fwd[r * L + j], b.inv[j * L + col]);
        acc = {acc.a + t.a, acc.b + t.b, acc.c + t.c, acc.d + t.d};
      }
      c.fwd[r * L + col] = acc;
    }
  return gate_ptm_sde(c) == 0;
}

// Operator-Schmidt rank over F_p (w -> 3^((p-1)/8)) of the gate's local matrix across the split of its qubits
// into those in cut mask A and the rest: rank of the realigned matrix R[(oa, ia)][(ob, ib)] = M[(oa, ob)][(ia, ib)].
int gate_cut_rank(const HostGate &g, int A) {
  const int L = 1 << g.nq;
  if ((int)g.fwd.size() != L * L) return -1;
  const uint64_t P = kPrime;
  auto mulm = [&](uint64_t a, uint64_t b) { return a * b % P; };
  auto powm = [&](uint64_t b, uint64_t e) { uint64_t r = 1; for (b %= P; e; e >>= 1, b = mulm(b, b)) if (e & 1) r = mulm(r, b); return r; };
  const uint64_t w = powm(3, (P - 1) / 8);
  const uint64_t wp[4] = {1, w, mulm(w, w), mulm(mulm(w, w), w)};
  auto img = [&](const Zq &z) {
    const int64_t cf[4] = {z.a, z.b, z.c, z.d};
    uint64_t s = 0;
    for (int i = 0; i < 4; i++) s = (s + mulm((uint64_t)(((cf[i] % (int64_t)P) + (int64_t)P) % (int64_t)P), wp[i])) % P;
    return s;
  };
  int la = 0;                                   // local qubit j of the gate is in A
  for (int j = 0; j < g.nq; j++) if ((A >> g.q[j]) & 1) la |= 1 << j;
  const int na = __builtin_popcount(la), nb = g.nq - na;
  if (!na || !nb) return 1;
  const int R = 1 << (2 * na), Cc = 1 << (2 * nb);
  std::vector<uint64_t> m((size_t)R * Cc, 0);
  auto split = [&](int x, int &xa, int &xb) {
    xa = xb = 0;
    int ia = 0, ib = 0;
    for (int j = 0; j < g.nq; j++) {
      const int bit = (x >> j) & 1;
      if ((la >> j) & 1) xa |= bit << ia++; else xb |= bit << ib++;
    }
  };
  for (int o = 0; o < L; o++)
    for (int i = 0; i < L; i++) {
      int oa, ob, ia, ib;
      split(o, oa, ob);
      split(i, ia, ib);
      m[(size_t)((oa << na) | ia) * Cc + ((ob << nb) | ib)] = img(g.fwd[o * L + i]);
    }
  int rank = 0;
  for (int col = 0; col < Cc && rank < R; col++) {
    int piv = -1;
    for (int r = rank; r < R; r++) if (m[(size_t)r * Cc + col]) { piv = r; break; }
    if (piv < 0) continue;
    for (int c = 0; c < Cc; c++) std::swap(m[(size_t)piv * Cc + c], m[(size_t)rank * Cc + c]);
    const uint64_t inv = powm(m[(size_t)rank * Cc + col], P - 2);
    for (int r = 0; r < R; r++) {
      if (r == rank || !m[(size_t)r * Cc + col]) continue;
      const uint64_t f = mulm(m[(size_t)r * Cc + col], inv);
      for (int c = 0; c < Cc; c++) m[(size_t)r * Cc + c] = (m[(size_t)r * Cc + c] + P - mulm(f, m[(size_t)rank * Cc + c])) % P;
    }
    rank++;
  }
  return rank;
}

}  // namespace chz
