// ORACLE (test infrastructure only): a scalar CPU restatement of the reference's exact gate-count
// branch-and-bound search, used as the parity checker for the HIP product path and as the timed CPU
// "port" baseline.  Only tests/, the smoke check and bench.py's cpu_baseline leg may load or run this.
// It shares no code with the product (cheese_amd/csrc).
//
// Followed reference semantics (file:line in the reference tree):
//   Z[w] ring, exact /sqrt2 .................... src/cheese.cu:59-104
//   residual update S <- S g^-1, normalise ...... src/cheese.cu:136-174
//   goal test ................................... src/cheese.cu:195-203
//   class key / exact class equality ........... src/cheese.cu:224-255 (own hash; the key only narrows)
//   F_p cut rank, support, H structure ......... src/cheese.cu:257-338
//   PTM sde (definition) and its WHT form ....... src/cheese.cu:340-449
//   CX bound, combine ........................... src/cheese.cu:585-607
//   window rules W2/W3 .......................... src/cheese.cu:609-632
//   per-child status (Cheap + Full passes) ...... src/cheese.cu:662-720
//   level loop, counts, dedupe, JSON line ....... src/cheese.cu:1249-1377
//   spec reader ................................. src/cheese.cu:1149-1226
//   symmetry group .............................. src/cheese.cu:1379-1429
//   run/main, depth-0 test, result line ......... src/cheese.cu:1658-1662, 1762-1895
// Parity pins: golden per-level logs of the reference CPU build (tests/golden/*.log), the published Toffoli
// D=11 frontier 49,503, the definitional PTM sde vs its WHT form, exact-vs-F_p ranks, closed-form gates.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

namespace orc {

constexpr int DMAX = 48, GMAX = 64, LMAX = 8, CUTMAX = 16;
constexpr uint64_t PR = 998244353ULL;

struct Z { int64_t a, b, c, d; };
static inline Z zadd(Z x, Z y) { return {x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; }
// (a + b w + c w^2 + d w^3)(a' + ...), w^4 = -1
static inline Z zmul(Z x, Z y) {
  return {x.a * y.a - x.b * y.d - x.c * y.c - x.d * y.b, x.a * y.b + x.b * y.a - x.c * y.d - x.d * y.c,
          x.a * y.c + x.b * y.b + x.c * y.a - x.d * y.d, x.a * y.d + x.b * y.c + x.c * y.b + x.d * y.a};
}
static inline Z zconj(Z x) { return {x.a, -x.d, -x.c, -x.b}; }
static inline bool zzero(Z x) { return !(x.a | x.b | x.c | x.d); }
static inline bool zeq(Z x, Z y) { return x.a == y.a && x.b == y.b && x.c == y.c && x.d == y.d; }
static inline bool zhalf(Z x) { return !((x.a - x.c) & 1) && !((x.b - x.d) & 1); }   // divisible by sqrt2
static inline Z zdown(Z x) { return {(x.b - x.d) / 2, (x.a + x.c) / 2, (x.b + x.d) / 2, (x.c - x.a) / 2}; }
static inline Z wpow(int p) {
  p &= 7;
  Z r{0, 0, 0, 0};
  const int64_t s = p < 4 ? 1 : -1;
  switch (p & 3) { case 0: r.a = s; break; case 1: r.b = s; break; case 2: r.c = s; break; default: r.d = s; }
  return r;
}
static inline int zval(Z x) {
  int v = 0;
  while (zhalf(x)) { x = zdown(x); if (++v > 400) break; }
  return v;
}
static inline bool fits32(int64_t v) { return v >= INT32_MIN && v <= INT32_MAX; }

struct Gate { int nq, q[3], ik; Z inv[LMAX * LMAX]; };

struct Spec {
  int n = 0, N = 0, depth = 0, standard = 0, lite = 0, tdiv = 1, G = 0, tk = 0, states = 0, gk = 0;
  std::vector<std::string> names;
  std::vector<Gate> gates;
  std::vector<uint8_t> bad2;
  std::vector<uint32_t> bad3;
  std::vector<Z> target, goal;
  std::vector<int> cutmask, cutsize, rad, cxtable;
  // symmetry group
  int nsym = 1;
  std::vector<uint8_t> bperm, conj;
  uint64_t w8 = 0;
};

// residual matrix: N*N entries over Z[w] / sqrt2^k, held in int64 but always within int32 (else ovf)
struct Mat { std::vector<Z> e; int k = 0; bool ovf = false; };

static uint64_t mulm(uint64_t a, uint64_t b) { return (unsigned __int128)a * b % PR; }
static uint64_t powm(uint64_t a, uint64_t e) { uint64_t r = 1; while (e) { if (e & 1) r = mulm(r, a); a = mulm(a, a); e >>= 1; } return r; }
static uint64_t tom(int64_t v) { int64_t m = v % (int64_t)PR; return (uint64_t)(m < 0 ? m + (int64_t)PR : m); }
static uint64_t zmodp(Z x, uint64_t w) {
  const uint64_t w2 = mulm(w, w), w3 = mulm(w2, w);
  return (tom(x.a) + mulm(tom(x.b), w) + mulm(tom(x.c), w2) + mulm(tom(x.d), w3)) % PR;
}

// S <- S g^-1 then lower k while every entry is divisible by sqrt2 (src/cheese.cu:136-174).  Overflow is
// judged on the un-normalised products, as the reference stores them before normalising.
static void apply_inv(Mat &S, const Gate &g, int N) {
  const int L = 1 << g.nq;
  int mask = 0;
  for (int j = 0; j < g.nq; j++) mask |= 1 << g.q[j];
  for (int base = 0; base < N; base++) {
    if (base & mask) continue;
    int idx[LMAX];
    for (int l = 0; l < L; l++) {
      int x = base;
      for (int j = 0; j < g.nq; j++) if ((l >> j) & 1) x |= 1 << g.q[j];
      idx[l] = x;
    }
    for (int r = 0; r < N; r++) {
      Z col[LMAX];
      for (int lc = 0; lc < L; lc++) {
        Z acc{0, 0, 0, 0};
        for (int m = 0; m < L; m++) {
          const Z gm = g.inv[m * L + lc];
          if (!zzero(gm)) acc = zadd(acc, zmul(S.e[r * N + idx[m]], gm));
        }
        col[lc] = acc;
      }
      for (int lc = 0; lc < L; lc++) {
        const Z v = col[lc];
        if (!fits32(v.a) || !fits32(v.b) || !fits32(v.c) || !fits32(v.d)) S.ovf = true;
        // the reference narrows to int32 here; an overflowing run aborts, so wrap like the narrowing does
        S.e[r * N + idx[lc]] = Z{(int32_t)v.a, (int32_t)v.b, (int32_t)v.c, (int32_t)v.d};
      }
    }
  }
  S.k += g.ik;
  while (S.k > 0) {
    bool all = true;
    for (const Z &x : S.e) if (!zhalf(x)) { all = false; break; }
    if (!all) break;
    for (Z &x : S.e) x = zdown(x);
    S.k--;
  }
}

static bool is_goal(const Mat &S, const Spec &sp) {
  if (S.k != sp.gk) return false;
  for (size_t i = 0; i < S.e.size(); i++) if (!zeq(S.e[i], sp.goal[i])) return false;
  return true;
}

// ---- bounds ----
// rank over F_p of the operator-Schmidt matrix across qubit set A (src/cheese.cu:270-303)
static int cut_rank(const Mat &S, int n, int A, int sa, uint64_t w) {
  const int N = 1 << n, R = 1 << (2 * sa), C = N * N / R;
  std::vector<uint64_t> M((size_t)R * C);
  for (int o = 0; o < N; o++)
    for (int i = 0; i < N; i++) {
      int oa = 0, ob = 0, ia = 0, ib = 0, na = 0, nb = 0;
      for (int q = 0; q < n; q++) {
        if ((A >> q) & 1) { oa |= ((o >> q) & 1) << na; ia |= ((i >> q) & 1) << na; na++; }
        else { ob |= ((o >> q) & 1) << nb; ib |= ((i >> q) & 1) << nb; nb++; }
      }
      M[(size_t)((oa << na) | ia) * C + ((ob << nb) | ib)] = zmodp(S.e[o * N + i], w);
    }
  int rank = 0;
  for (int c = 0; c < C && rank < R; c++) {
    int piv = -1;
    for (int r = rank; r < R; r++) if (M[(size_t)r * C + c]) { piv = r; break; }
    if (piv < 0) continue;
    if (piv != rank) for (int j = 0; j < C; j++) std::swap(M[(size_t)piv * C + j], M[(size_t)rank * C + j]);
    const uint64_t inv = powm(M[(size_t)rank * C + c], PR - 2);
    for (int r = rank + 1; r < R; r++) {
      const uint64_t x = M[(size_t)r * C + c];
      if (!x) continue;
      const uint64_t f = mulm(x, inv);
      for (int j = c; j < C; j++) M[(size_t)r * C + j] = (M[(size_t)r * C + j] + PR - mulm(f, M[(size_t)rank * C + j])) % PR;
    }
    rank++;
  }
  return rank;
}

// number of qubits S acts on non-trivially (src/cheese.cu:305-319)
static int support(const Mat &S, int n) {
  const int N = 1 << n;
  int cnt = 0;
  for (int q = 0; q < n; q++) {
    const int bit = 1 << q;
    bool triv = true;
    for (int r = 0; r < N && triv; r++)
      for (int c = 0; c < N && triv; c++) {
        if ((r ^ c) & bit) { if (!zzero(S.e[r * N + c])) triv = false; }
        else if (!((r | c) & bit) && !zeq(S.e[r * N + c], S.e[(r | bit) * N + (c | bit)])) triv = false;
      }
    cnt += !triv;
  }
  return cnt;
}

// 2 when S is monomial with a permutation that is not affine over GF(2) (src/cheese.cu:321-338)
static int h_struct(const Mat &S, int n) {
  const int N = 1 << n;
  int pi[16];
  for (int c = 0; c < N; c++) {
    int nz = 0, row = 0;
    for (int r = 0; r < N; r++) if (!zzero(S.e[r * N + c])) { nz++; row = r; }
    if (nz != 1) return 0;
    pi[c] = row;
  }
  for (int j = 0; j < N; j++) {
    int x = pi[0];
    for (int i = 0; i < n; i++) if ((j >> i) & 1) x ^= pi[1 << i] ^ pi[0];
    if (pi[j] != x) return 2;
  }
  return 0;
}

// Pauli digit per qubit 0 I, 1 X, 2 Y, 3 Z (src/cheese.cu:343-355)
static void pauli(int b, int nq, int &flip, int &ym, int &zm) {
  flip = ym = zm = 0;
  for (int q = 0; q < nq; q++) {
    const int d = (b >> (2 * q)) & 3;
    if (d == 1 || d == 2) flip |= 1 << q;
    if (d == 2) ym |= 1 << q;
    if (d == 3) zm |= 1 << q;
  }
}
static int pphase(int x, int ym, int zm) { return (2 * __builtin_popcount(ym) + 4 * __builtin_popcount(x & (ym | zm))) & 7; }

// definitional T lower bound: max_ab 2(n+k) - v(Tr(P_a S' P_b S'^dag)) (src/cheese.cu:357-395)
static int ptm_sde(const Mat &S, int n) {
  const int N = 1 << n, P = 1 << (2 * n);
  if (S.k + 2 * n + 6 > 62) return 0;
  int best = 0;
  std::vector<Z> SP(N * N), M(N * N);
  for (int b = 0; b < P; b++) {
    int fb, yb, zb;
    pauli(b, n, fb, yb, zb);
    for (int r = 0; r < N; r++)
      for (int c = 0; c < N; c++) {
        const Z x = S.e[r * N + (c ^ fb)];
        SP[r * N + c] = zzero(x) ? x : zmul(x, wpow(pphase(c, yb, zb)));
      }
    for (int r = 0; r < N; r++)
      for (int s = 0; s < N; s++) {
        Z acc{0, 0, 0, 0};
        for (int c = 0; c < N; c++) acc = zadd(acc, zmul(SP[r * N + c], zconj(S.e[s * N + c])));
        M[r * N + s] = acc;
      }
    for (int a = 0; a < P; a++) {
      int fa, ya, za;
      pauli(a, n, fa, ya, za);
      Z acc{0, 0, 0, 0};
      for (int x = 0; x < N; x++) acc = zadd(acc, zmul(M[(x ^ fa) * N + x], wpow(pphase(x ^ fa, ya, za))));
      if (!zzero(acc)) best = std::max(best, 2 * (n + S.k) - zval(acc));
    }
  }
  return best;
}

// same bound by 2-D Walsh-Hadamard transforms over Z[w] mod 2^b (src/cheese.cu:397-449)
template <typename U> static int ptm_wht_t(const Mat &S, int n) {
  const int N = 1 << n, E = N * N, top = 2 * (n + S.k);
  int best = 0;
  std::vector<U> qa(E), qb(E), qc(E), qd(E);
  for (int fa = 0; fa < N; fa++)
    for (int fb = 0; fb < N; fb++) {
      for (int e = 0; e < E; e++) {
        const int x = e >> n, c = e & (N - 1);
        const Z u = S.e[(x ^ fa) * N + (c ^ fb)], y = zconj(S.e[e]);
        // wrapping products in U (the low bits are all the valuation needs)
        const U xa = (U)u.a, xb = (U)u.b, xc = (U)u.c, xd = (U)u.d, ya = (U)y.a, yb = (U)y.b, yc = (U)y.c, yd = (U)y.d;
        qa[e] = U(xa * ya - xb * yd - xc * yc - xd * yb);
        qb[e] = U(xa * yb + xb * ya - xc * yd - xd * yc);
        qc[e] = U(xa * yc + xb * yb + xc * ya - xd * yd);
        qd[e] = U(xa * yd + xb * yc + xc * yb + xd * ya);
      }
      for (int h = 1; h < E; h <<= 1)
        for (int e = 0; e < E; e++)
          if (!(e & h)) {
            const int f = e | h;
            U t;
            t = qa[e]; qa[e] = U(t + qa[f]); qa[f] = U(t - qa[f]);
            t = qb[e]; qb[e] = U(t + qb[f]); qb[f] = U(t - qb[f]);
            t = qc[e]; qc[e] = U(t + qc[f]); qc[f] = U(t - qc[f]);
            t = qd[e]; qd[e] = U(t + qd[f]); qd[f] = U(t - qd[f]);
          }
      for (int e = 0; e < E; e++) {
        const U o = qa[e] | qb[e] | qc[e] | qd[e];
        if (!o) continue;
        const int m = __builtin_ctzll((unsigned long long)o);
        const int v = 2 * m + (!(((qa[e] ^ qc[e]) >> m) & 1) && !(((qb[e] ^ qd[e]) >> m) & 1));
        best = std::max(best, top - v);
      }
    }
  return best;
}
// note: products of two int32-range entries can exceed int64 only beyond 2^62; stored entries are int32
static int ptm_wht(const Mat &S, int n) {
  return 2 * (n + S.k) <= 64 ? ptm_wht_t<uint32_t>(S, n) : ptm_wht_t<uint64_t>(S, n);
}

static int cx_bound(const Mat &S, const Spec &sp) {
  int idx = 0, mult = 1, mx = 0;
  for (size_t i = 0; i < sp.cutmask.size(); i++) {
    const int rk = cut_rank(S, sp.n, sp.cutmask[i], sp.cutsize[i], sp.w8);
    int need = 0;
    while ((1 << need) < rk) need++;
    need = std::min(need, sp.rad[i] - 1);
    idx += need * mult;
    mult *= sp.rad[i];
    mx = std::max(mx, need);
  }
  return sp.cxtable.empty() ? mx : sp.cxtable[idx];
}

static int combine(int h, int t, int c2, int supp) {
  int best = 1 << 30;
  for (int c = c2; c <= std::max(c2, supp); c++) best = std::min(best, c + std::max(h + t, supp - 2 * c));
  return best;
}

// ---- class key and exact class equality ----
static inline uint64_t smix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}
static uint64_t mat_hash(const Mat &S, const Spec &sp, int s) {
  const int N = sp.N, n = sp.n;
  const uint8_t *pm = &sp.bperm[(size_t)s * N];
  uint64_t acc = 0;
  for (int e = 0; e < N * N; e++) {
    Z v = S.e[e];
    if (sp.conj[s]) v = zconj(v);
    const int pos = pm[e >> n] * N + pm[e & (N - 1)];
    uint64_t h = smix((uint64_t)pos * 0x100000001b3ULL ^ (uint64_t)(uint32_t)v.a);
    h = smix(h ^ ((uint64_t)(uint32_t)v.b << 32 | (uint32_t)v.c));
    h = smix(h ^ (uint64_t)(uint32_t)v.d);
    acc += h;
  }
  return smix(acc ^ (uint64_t)S.k * 0x51ed27ULL);
}
static uint64_t class_key(const Mat &S, const Spec &sp) {
  uint64_t best = ~0ULL;
  for (int s = 0; s < sp.nsym; s++) best = std::min(best, mat_hash(S, sp, s));
  return best;
}
static bool same_class(const Mat &A, const Mat &B, const Spec &sp) {
  const int N = sp.N, n = sp.n;
  if (A.k != B.k) return false;
  for (int s = 0; s < sp.nsym; s++) {
    const uint8_t *pm = &sp.bperm[(size_t)s * N];
    bool eq = true;
    for (int e = 0; eq && e < N * N; e++) {
      const Z v = sp.conj[s] ? zconj(A.e[e]) : A.e[e];
      eq = zeq(v, B.e[pm[e >> n] * N + pm[e & (N - 1)]]);
    }
    if (eq) return true;
  }
  return false;
}

// ---- spec reader (src/cheese.cu:1149-1226); returns 0 or 2 (bad spec) ----
static int read_spec(const std::string &text, Spec &s, std::string &err) {
  std::istringstream in(text);
  std::string tok;
  in >> tok;
  if (tok != "cheese-spec") { err = "not a spec"; return 2; }
  int ver; in >> ver;
  std::vector<Z> start;
  while (in >> tok) {
    if (tok == "n") in >> s.n;
    else if (tok == "mode" || tok == "phase") {
      std::string v; in >> v;
      if (v == "states") s.states = 1;
      else if (v != "unitary" && v != "exact") { err = "unsupported " + tok + " " + v; return 2; }
    } else if (tok == "pairs") { int m; in >> m; }
    else if (tok == "start") { int k; in >> k; const int N = 1 << s.n; start.resize(N * N); for (auto &x : start) in >> x.a >> x.b >> x.c >> x.d; }
    else if (tok == "goal") { int m; in >> m; int w = 2 + (int)s.cutmask.size(); for (int i = 0; i < m * w; i++) { int v; in >> v; } }
    else if (tok == "depth") in >> s.depth;
    else if (tok == "standard") in >> s.standard;
    else if (tok == "lite") in >> s.lite;
    else if (tok == "tdiv") in >> s.tdiv;
    else if (tok == "gates") {
      in >> s.G;
      if (s.G > GMAX) { err = "too many gates"; return 2; }
      for (int g = 0; g < s.G; g++) {
        std::string name; Gate gt{}; int gk;
        in >> name >> gt.nq;
        for (int j = 0; j < gt.nq; j++) in >> gt.q[j];
        const int L = 1 << gt.nq;
        in >> gk;
        Z fwd[LMAX * LMAX];
        for (int i = 0; i < L * L; i++) in >> fwd[i].a >> fwd[i].b >> fwd[i].c >> fwd[i].d;
        in >> gt.ik;
        for (int i = 0; i < L * L; i++) in >> gt.inv[i].a >> gt.inv[i].b >> gt.inv[i].c >> gt.inv[i].d;
        s.names.push_back(name); s.gates.push_back(gt);
      }
    } else if (tok == "bad2") {
      int m; in >> m; s.bad2.assign((size_t)s.G * s.G, 0);
      for (int i = 0; i < m; i++) { int a, b; in >> a >> b; s.bad2[a * s.G + b] = 1; }
    } else if (tok == "bad3") {
      int m; in >> m; s.bad3.assign(((size_t)s.G * s.G * s.G + 31) / 32, 0);
      for (int i = 0; i < m; i++) { int a, b, c; in >> a >> b >> c; const uint32_t t = (a * s.G + b) * s.G + c; s.bad3[t >> 5] |= 1u << (t & 31); }
    } else if (tok == "target") {
      int k; in >> k; const int N = 1 << s.n; std::vector<Z> M(N * N);
      for (auto &x : M) in >> x.a >> x.b >> x.c >> x.d;
      if (s.states) { s.goal = M; s.gk = k; }
      else { s.target = M; s.tk = k; s.gk = 0; s.goal.assign(N * N, Z{0, 0, 0, 0}); for (int i = 0; i < N; i++) s.goal[i * N + i] = Z{1, 0, 0, 0}; }
    } else if (tok == "cuts") {
      int m; in >> m;
      for (int i = 0; i < m; i++) {
        int sz, mask = 0; in >> sz;
        for (int j = 0; j < sz; j++) { int q; in >> q; mask |= 1 << q; }
        s.cutmask.push_back(mask); s.cutsize.push_back(sz); s.rad.push_back(2 * sz + 1);
      }
    } else if (tok == "cxtable") { int m; in >> m; s.cxtable.resize(m); for (auto &v : s.cxtable) in >> v; }
    else if (tok == "end") break;
    else { err = "unknown spec field " + tok; return 2; }
  }
  if (s.bad2.empty()) s.bad2.assign((size_t)s.G * s.G, 0);
  if (s.bad3.empty()) s.bad3.assign(((size_t)s.G * s.G * s.G + 31) / 32, 0);
  s.N = 1 << s.n;
  if (s.tdiv < 1) s.tdiv = 1;
  s.w8 = powm(3, (PR - 1) / 8);
  return 0;
}

// ---- symmetry group (src/cheese.cu:1379-1429) ----
static std::vector<std::string> split_us(const std::string &s) {
  std::vector<std::string> out;
  size_t a = 0, b;
  while ((b = s.find('_', a)) != std::string::npos) { out.push_back(s.substr(a, b - a)); a = b + 1; }
  out.push_back(s.substr(a));
  return out;
}
static int map_gate(const Spec &s, const std::vector<int> &perm, bool cj, int g) {
  const std::vector<std::string> p = split_us(s.names[g]);
  if (p.size() < 2) return -1;
  std::string gate = p[1];
  if (cj) {
    if (gate == "t") gate = "tdg"; else if (gate == "tdg") gate = "t";
    else if (gate == "s") gate = "sdg"; else if (gate == "sdg") gate = "s";
    else if (gate != "x" && gate != "z" && gate != "h" && gate != "cx" && gate != "cz" && gate != "ccx") return -1;
  }
  std::vector<int> qs;
  for (size_t i = 2; i < p.size(); i++) {
    const int q = std::atoi(p[i].c_str());
    if (q < 0 || q >= (int)perm.size()) return -1;
    qs.push_back(perm[q]);
  }
  if (gate == "ccx" && qs.size() == 3 && qs[0] > qs[1]) std::swap(qs[0], qs[1]);
  std::string want = "L_" + gate;
  for (int q : qs) want += "_" + std::to_string(q);
  for (int h = 0; h < s.G; h++) if (s.names[h] == want) return h;
  return -1;
}
static void build_sym(Spec &s, bool on) {
  const int n = s.n, N = 1 << n;
  std::vector<int> p(n);
  for (int i = 0; i < n; i++) p[i] = i;
  std::vector<std::vector<int>> perms;
  do perms.push_back(p); while (on && std::next_permutation(p.begin(), p.end()));
  s.bperm.clear(); s.conj.clear();
  for (int cj = 0; cj < (on ? 2 : 1); cj++)
    for (auto &pm : perms) {
      bool ok = true;
      for (int g = 0; g < s.G && ok; g++) ok = map_gate(s, pm, cj, g) >= 0;
      if (!ok) continue;
      s.conj.push_back((uint8_t)cj);
      for (int x = 0; x < N; x++) {
        int y = 0;
        for (int i = 0; i < n; i++) if ((x >> i) & 1) y |= 1 << pm[i];
        s.bperm.push_back((uint8_t)y);
      }
    }
  s.nsym = (int)s.conj.size();
}

enum : uint8_t { DEAD = 0, ALIVE = 1, FOUND = 2, W2 = 3, W3 = 4, OVF = 6 };

struct Stats { long cheap = 0, tbound = 0, hashed = 0; };

struct Node { uint32_t parent; uint8_t gate; };

static double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// the whole run; returns the reference exit code, appends stdout lines to out
static int run(Spec &sp, int depth, bool dedupe, bool sym, std::string &out, std::string &err, Stats *st) {
  char buf[512];
  if (sp.states) { err = "state targets are not part of this build"; return 3; }
  if (sp.n < 1 || sp.n > 3) { err = "n not supported"; return 3; }
  if (depth > DMAX) { err = "depth > 48"; return 3; }
  build_sym(sp, sym);
  const int N = sp.N, E = N * N, G = sp.G;
  Mat U; U.e.resize(E); U.k = sp.tk;
  for (int i = 0; i < E; i++) {
    const Z x = sp.target[i];
    if (!fits32(x.a) || !fits32(x.b) || !fits32(x.c) || !fits32(x.d)) U.ovf = true;
    U.e[i] = x;
  }
  if (U.ovf) { err = "target coefficients exceed int32"; return 4; }
  if (is_goal(U, sp)) { out += "{\"found\": 0, \"circuit\": []}\n"; return 0; }
  const double t0 = now();
  std::vector<std::vector<Node>> tree(1);        // level 0: the empty prefix
  std::vector<Mat> front{U};
  std::vector<int> g1s{-1}, g2s{-1};
  for (int t = 1; t <= depth; t++) {
    const double tl = now();
    const long m = (long)front.size();
    const bool keep = t < depth;
    const long nc = m * G;
    std::vector<uint8_t> status(nc);
    std::vector<uint64_t> key(nc);
    long ccheap = 0, ct = 0, ch = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : ccheap, ct, ch)
    for (long o = 0; o < nc; o++) {
      const long i = o / G;
      const int g = (int)(o % G), g1 = g1s[i], g2 = g2s[i];
      if (g1 >= 0 && sp.bad2[g1 * G + g]) { status[o] = W2; continue; }
      if (g2 >= 0) {
        const uint32_t tri = ((uint32_t)g2 * G + g1) * G + g;
        if ((sp.bad3[tri >> 5] >> (tri & 31)) & 1) { status[o] = W3; continue; }
      }
      Mat C = front[i];
      apply_inv(C, sp.gates[g], N);
      if (C.ovf) { status[o] = OVF; continue; }
      if (is_goal(C, sp)) { status[o] = FOUND; continue; }
      if (sp.standard || sp.lite) {
        ccheap++;
        int h, supp = 0, c2 = 0;
        if (sp.lite) h = C.k;
        else { h = std::max(C.k, h_struct(C, sp.n)); supp = support(C, sp.n); c2 = cx_bound(C, sp); }
        if (combine(h, 0, c2, supp) > depth - t) { status[o] = DEAD; continue; }
        ct++;
        const int T = (ptm_wht(C, sp.n) + sp.tdiv - 1) / sp.tdiv;
        if (combine(h, T, c2, supp) > depth - t) { status[o] = DEAD; continue; }
      }
      status[o] = ALIVE;
      if (keep) { key[o] = class_key(C, sp); ch++; }
    }
    if (st) { st->cheap += ccheap; st->tbound += ct; st->hashed += ch; }
    long n_w2 = 0, n_w3 = 0, n_dead = 0, n_found = 0, n_alive = 0, first_found = -1;
    for (long o = 0; o < nc; o++) {
      switch (status[o]) {
        case W2: n_w2++; break;
        case W3: n_w3++; break;
        case DEAD: n_dead++; break;
        case FOUND: n_found++; if (first_found < 0) first_found = o; break;
        case ALIVE: n_alive++; break;
        case OVF: err = "residual coefficients exceed the storage type"; return 4;
      }
    }
    if (n_found) {
      std::vector<int> seq(t);
      seq[t - 1] = (int)(first_found % G);
      uint32_t i = (uint32_t)(first_found / G);
      for (int l = t - 1; l >= 1; l--) { seq[l - 1] = tree[l][i].gate; i = tree[l][i].parent; }
      std::string s = "{\"t\": " + std::to_string(t) + ", \"found\": " + std::to_string(n_found) + ", \"circuit\": [";
      for (int j = 0; j < t; j++) s += (j ? ", \"" : "\"") + sp.names[seq[j]] + "\"";
      std::snprintf(buf, sizeof buf, "], \"seconds\": %.2f}\n", now() - t0);
      out += s + buf;
      return 0;
    }
    // survivors in generation (= lexicographic) order
    std::vector<long> surv;
    if (keep) { surv.reserve(n_alive); for (long o = 0; o < nc; o++) if (status[o] == ALIVE) surv.push_back(o); }
    std::vector<char> kept(surv.size(), 1);
    if (dedupe && surv.size() > 1) {
      std::vector<long> ord(surv.size());
      for (size_t j = 0; j < ord.size(); j++) ord[j] = (long)j;
      std::stable_sort(ord.begin(), ord.end(), [&](long a, long b) { return key[surv[a]] < key[surv[b]]; });
      // runs of equal key: a member is dropped when it is in the class of an earlier kept member
      std::vector<std::pair<size_t, size_t>> runs;
      for (size_t a = 0; a < ord.size();) {
        size_t b = a + 1;
        while (b < ord.size() && key[surv[ord[b]]] == key[surv[ord[a]]]) b++;
        if (b - a > 1) runs.push_back({a, b});
        a = b;
      }
#pragma omp parallel for schedule(dynamic, 16)
      for (size_t r = 0; r < runs.size(); r++) {
        std::vector<Mat> reps;
        for (size_t j = runs[r].first; j < runs[r].second; j++) {
          const long o = surv[ord[j]];
          Mat C = front[o / G];
          apply_inv(C, sp.gates[o % G], N);
          bool dup = false;
          for (const Mat &R : reps) if (same_class(C, R, sp)) { dup = true; break; }
          if (dup) kept[ord[j]] = 0; else reps.push_back(C);
        }
      }
    }
    long distinct = 0;
    for (char k : kept) distinct += k;
    if (!keep) distinct = 0;
    std::snprintf(buf, sizeof buf,
                  "{\"t\": %d, \"parents\": %ld, \"prefixes\": %ld, \"cut_w3\": %ld, \"cut_lb\": %ld, \"survivors\": %ld, "
                  "\"distinct\": %ld, \"proved_unsat\": %s, \"seconds\": %.2f}\n",
                  t, m, m * G - n_w2, n_w3, n_dead, n_alive, distinct, distinct ? "false" : "true", now() - tl);
    out += buf;
    if (!distinct) return 1;
    std::vector<Node> lvl;
    std::vector<Mat> nf;
    std::vector<int> n1, n2;
    lvl.reserve(distinct); nf.reserve(distinct);
    for (size_t j = 0; j < surv.size(); j++) {
      if (!kept[j]) continue;
      const long o = surv[j];
      const uint32_t par = (uint32_t)(o / G);
      const int g = (int)(o % G);
      lvl.push_back({par, (uint8_t)g});
      Mat C = front[par];
      apply_inv(C, sp.gates[g], N);
      nf.push_back(std::move(C));
      n1.push_back(g);
      n2.push_back(g1s[par]);
    }
    tree.push_back(std::move(lvl));
    front.swap(nf); g1s.swap(n1); g2s.swap(n2);
  }
  return 1;   // depth 0 with a non-identity target: nothing to expand
}

}  // namespace orc

// ---------------- C ABI used by the tests and the CPU-baseline leg ----------------
extern "C" {

// Runs the search on spec text; writes the stdout lines (NUL-terminated, truncated to cap) to out.
// depth < 0: the spec's depth.  Returns the reference exit code (0 found, 1 proof, 2 bad spec, 3 usage,
// 4 overflow).  stats (3 longs, may be NULL): children through the cheap bounds, the T bound, the hash.
int orc_run(const char *spec, long len, int depth, int dedupe, int sym, int threads, char *out, long cap, long *stats) {
  if (threads > 0) omp_set_num_threads(threads);
  orc::Spec sp;
  std::string err, o;
  int rc = orc::read_spec(std::string(spec, (size_t)len), sp, err);
  if (rc == 0) {
    const int d = depth >= 0 ? depth : sp.depth;
    orc::Stats st;
    rc = orc::run(sp, d, dedupe != 0, sym != 0, o, err, &st);
    if (stats) { stats[0] = st.cheap; stats[1] = st.tbound; stats[2] = st.hashed; }
  }
  if (!err.empty()) std::fprintf(stderr, "oracle: %s\n", err.c_str());
  if (out && cap > 0) { const size_t n = std::min((size_t)cap - 1, o.size()); std::memcpy(out, o.data(), n); out[n] = 0; }
  return rc;
}

// Exact residual of a gate sequence applied to the spec's target (S = U g_1^-1 ... g_t^-1), entries as
// (a,b,c,d) int64 into e[4*N*N]; returns k, or -1 on a bad spec / overflow.
int orc_residual(const char *spec, long len, const int *seq, int t, long long *e) {
  orc::Spec sp;
  std::string err;
  if (orc::read_spec(std::string(spec, (size_t)len), sp, err)) return -1;
  orc::Mat S; S.e = sp.target; S.k = sp.tk;
  for (int i = 0; i < t; i++) { if (seq[i] < 0 || seq[i] >= sp.G) return -1; orc::apply_inv(S, sp.gates[seq[i]], sp.N); }
  if (S.ovf) return -1;
  for (size_t i = 0; i < S.e.size(); i++) { e[4 * i] = S.e[i].a; e[4 * i + 1] = S.e[i].b; e[4 * i + 2] = S.e[i].c; e[4 * i + 3] = S.e[i].d; }
  return S.k;
}

// Single-matrix bounds for unit parity tests: entries e[4*N*N] (int64 coefficients, int32 range), exponent k.
// which: 0 T bound by WHT, 1 T bound by definition, 2 support, 3 H structure, 4 F_p rank across cut mask A.
int orc_bound(int n, int k, const long long *e, int which, int A) {
  orc::Mat S; S.k = k; S.e.resize((size_t)1 << (2 * n));
  for (size_t i = 0; i < S.e.size(); i++) S.e[i] = {e[4 * i], e[4 * i + 1], e[4 * i + 2], e[4 * i + 3]};
  switch (which) {
    case 0: return orc::ptm_wht(S, n);
    case 1: return orc::ptm_sde(S, n);
    case 2: return orc::support(S, n);
    case 3: return orc::h_struct(S, n);
    case 4: return orc::cut_rank(S, n, A, __builtin_popcount(A), orc::powm(3, (orc::PR - 1) / 8));
  }
  return -1;
}

}  // extern "C"
