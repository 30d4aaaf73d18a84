// T-count host side: the .tspec text of a unitary (its Pauli transfer matrix and the rotation tables the search
// peels with) and the exact check of a T-count decomposition.  Host only, no device work.
//
// Paulis are numbered p = sum_q d_q 4^q with d = 0 I, 1 X, 2 Y, 3 Z on qubit q (the .tspec convention of
// python/cheese/tcount.py:20-36).  Here a Pauli is handled in its symplectic form: an X-part bit mask (d = X or Y)
// and a Z-part bit mask (d = Y or Z); products and commutation come from those masks, not from matrices.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cheese_hip.h"
#include "spec.h"

namespace {

using chz::Zq;

Zq zmul(const Zq &x, const Zq &y) {
  return {x.a * y.a - x.b * y.d - x.c * y.c - x.d * y.b, x.a * y.b + x.b * y.a - x.c * y.d - x.d * y.c,
          x.a * y.c + x.b * y.b + x.c * y.a - x.d * y.d, x.a * y.d + x.b * y.c + x.c * y.b + x.d * y.a};
}
Zq zadd(const Zq &x, const Zq &y) { return {x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; }
Zq zconj(const Zq &x) { return {x.a, -x.d, -x.c, -x.b}; }   // w -> w^7 = -w^3
// w^p times x: multiplying by w shifts the coefficients up one place, the top one wrapping with a sign (w^4 = -1)
Zq zrot(Zq x, int p) {
  for (p &= 7; p > 0; p--) x = {-x.d, x.a, x.b, x.c};
  return x;
}
bool zzero(const Zq &x) { return !(x.a | x.b | x.c | x.d); }

// a Pauli operator in symplectic form
struct Pauli { uint32_t x, z; };

Pauli pauli_of(int n, int p) {
  Pauli P{0, 0};
  for (int q = 0; q < n; q++) {
    const int d = (p >> (2 * q)) & 3;
    if (d == 1 || d == 2) P.x |= 1u << q;
    if (d == 2 || d == 3) P.z |= 1u << q;
  }
  return P;
}

int index_of(int n, Pauli P) {
  static const int digit[2][2] = {{0, 3}, {1, 2}};   // [x bit][z bit]
  int p = 0;
  for (int q = 0; q < n; q++) p |= digit[(P.x >> q) & 1][(P.z >> q) & 1] << (2 * q);
  return p;
}

// P|b> = w^phase(b) |b ^ P.x>: per qubit, Z contributes (-1)^bit and Y = iXZ contributes i (-1)^bit
int pauli_phase(Pauli P, uint32_t b) {
  const int y = __builtin_popcount(P.x & P.z), zb = __builtin_popcount(P.z & b);
  return 2 * y + 4 * zb;
}

// P_a P_c = i^m P_{a xor c}: the i-power m (mod 4) of the product of two Paulis, from the phases of their
// actions on a basis state (the phase of Y Y... appears once per Y in each factor)
int product_ipower(Pauli A, Pauli C) {
  // (P_a P_c)|0> = P_a (w^{phase_c(0)} |C.x>) = w^{phase_c(0) + phase_a(C.x)} |A.x ^ C.x>, and the product Pauli
  // R = A xor C sends |0> to w^{phase_r(0)} |R.x>; the ratio is i^m
  const Pauli R{A.x ^ C.x, A.z ^ C.z};
  const int w8 = pauli_phase(C, 0) + pauli_phase(A, C.x) - pauli_phase(R, 0);
  return ((w8 % 8 + 8) % 8) / 2;
}

bool anticommute(Pauli A, Pauli C) {
  return (__builtin_popcount(A.x & C.z) + __builtin_popcount(A.z & C.x)) & 1;
}

struct Mat {
  int N = 0, k = 0;
  std::vector<Zq> e;   // row-major
};

// divide by sqrt2 while every entry allows it (the normal form: equal matrices, equal representation)
void normalise(Mat &M) {
  for (;;) {
    if (M.k <= 0) return;
    for (const Zq &x : M.e)
      if (((x.a - x.c) & 1) || ((x.b - x.d) & 1)) return;
    for (Zq &x : M.e) x = {(x.b - x.d) / 2, (x.a + x.c) / 2, (x.b + x.d) / 2, (x.c - x.a) / 2};
    M.k--;
  }
}

Mat matmul(const Mat &A, const Mat &B) {
  Mat C;
  C.N = A.N;
  C.k = A.k + B.k;
  C.e.assign((size_t)A.N * A.N, Zq{0, 0, 0, 0});
  for (int i = 0; i < A.N; i++)
    for (int l = 0; l < A.N; l++) {
      const Zq a = A.e[i * A.N + l];
      if (zzero(a)) continue;
      for (int j = 0; j < A.N; j++) C.e[i * A.N + j] = zadd(C.e[i * A.N + j], zmul(a, B.e[l * A.N + j]));
    }
  normalise(C);
  return C;
}

// the transfer matrix Uc[a][b] = Tr(P_a U P_b U^dag) / 2^n as (x + y sqrt2) / sqrt2^e with one exponent e,
// normalised.  Tr(P_a U P_b U^dag) = sum_{j,l} w^{phase_a(j) + phase_b(l)} U[j][l ^ b.x] conj(U[j ^ a.x][l]):
// both Paulis are signed permutations, so the trace is a double sum over basis states.  The entries of U carry
// 1/sqrt2^k, so every trace carries 1/sqrt2^(2k), and 1/2^n = 1/sqrt2^(2n).
// Returns false when a trace is not real (U not unitary) or leaves int64.
bool transfer_matrix(int n, const Mat &U, int &e, std::vector<std::pair<long long, long long>> &xy) {
  const int N = 1 << n, P = 1 << (2 * n);
  std::vector<Pauli> paulis(P);
  for (int p = 0; p < P; p++) paulis[p] = pauli_of(n, p);
  xy.assign((size_t)P * P, {0, 0});
  for (int b = 0; b < P; b++) {
    const Pauli B = paulis[b];
    for (int a = 0; a < P; a++) {
      const Pauli A = paulis[a];
      Zq tr{0, 0, 0, 0};
      for (int j = 0; j < N; j++) {
        const int pa = pauli_phase(A, (uint32_t)j);
        const int ja = j ^ (int)A.x;
        for (int l = 0; l < N; l++) {
          const Zq u = U.e[j * N + (l ^ (int)B.x)];
          if (zzero(u)) continue;
          const Zq v = U.e[ja * N + l];
          if (zzero(v)) continue;
          tr = zadd(tr, zrot(zmul(u, zconj(v)), pa + pauli_phase(B, (uint32_t)l)));
        }
      }
      // real elements of Z[w] are a + b sqrt2 = a + b (w - w^3): c = 0 and d = -b
      if (tr.c != 0 || tr.d != -tr.b) return false;
      xy[(size_t)a * P + b] = {tr.a, tr.b};
    }
  }
  e = 2 * (n + U.k);
  // (x + y sqrt2) / sqrt2 = y + (x / 2) sqrt2: one exponent down while every x is even
  for (;;) {
    if (e <= 0) break;
    bool even = true;
    for (const auto &v : xy) even = even && !(v.first & 1);
    if (!even) break;
    for (auto &v : xy) v = {v.second, v.first / 2};
    e--;
  }
  return true;
}

bool load_unitary(int n, int k, const long long *entries, Mat &U) {
  if (n < 1 || n > 3 || k < 0 || !entries) return false;
  U.N = 1 << n;
  U.k = k;
  U.e.resize((size_t)U.N * U.N);
  for (size_t i = 0; i < U.e.size(); i++)
    U.e[i] = {entries[4 * i], entries[4 * i + 1], entries[4 * i + 2], entries[4 * i + 3]};
  return true;
}

// R(P_p) = ((1 + w) I + (1 - w) P_p) / 2, or its inverse ((1 + w^7) I + (1 - w^7) P_p) / 2
Mat rotation(int n, int p, bool inverse) {
  const int N = 1 << n;
  const Pauli Pp = pauli_of(n, p);
  const Zq one_plus = inverse ? Zq{1, 0, 0, -1} : Zq{1, 1, 0, 0};     // 1 + w or 1 + w^7 = 1 - w^3
  const Zq one_minus = inverse ? Zq{1, 0, 0, 1} : Zq{1, -1, 0, 0};
  Mat R;
  R.N = N;
  R.k = 2;
  R.e.assign((size_t)N * N, Zq{0, 0, 0, 0});
  for (int i = 0; i < N; i++) R.e[i * N + i] = one_plus;
  for (int b = 0; b < N; b++) {   // column b of P_p: w^phase(b) at row b ^ x
    const int r = b ^ (int)Pp.x;
    R.e[r * N + b] = zadd(R.e[r * N + b], zrot(one_minus, pauli_phase(Pp, (uint32_t)b)));
  }
  normalise(R);
  return R;
}

}  // namespace

extern "C" {

int chz_tspec_write(int n, int k, const long long *entries, char *out, size_t cap) {
  Mat U;
  if (!load_unitary(n, k, entries, U)) return -3;
  int e = 0;
  std::vector<std::pair<long long, long long>> xy;
  if (!transfer_matrix(n, U, e, xy)) return -2;
  const int P = 1 << (2 * n);
  std::string s = "tcount-spec 1\nn " + std::to_string(n) + "\nchannel " + std::to_string(e) + "\n";
  char buf[64];
  for (int a = 0; a < P; a++) {
    for (int b = 0; b < P; b++) {
      std::snprintf(buf, sizeof buf, b ? " %lld %lld" : "%lld %lld", xy[(size_t)a * P + b].first,
                    xy[(size_t)a * P + b].second);
      s += buf;
    }
    s += "\n";
  }
  // rotation P_p (p >= 1) against each Pauli P_c: "-1 0" when they commute, else (r, s) with i P_p P_c = s P_r
  s += "rotations " + std::to_string(P - 1) + "\n";
  for (int p = 1; p < P; p++) {
    const Pauli A = pauli_of(n, p);
    for (int c = 0; c < P; c++) {
      const Pauli C = pauli_of(n, c);
      int r = -1, sign = 0;
      if (anticommute(A, C)) {
        r = index_of(n, Pauli{A.x ^ C.x, A.z ^ C.z});
        // i P_p P_c = i^(m+1) P_r with m odd: +1 when m + 1 = 0 mod 4, -1 when m + 1 = 2 mod 4
        sign = ((product_ipower(A, C) + 1) & 3) == 0 ? 1 : -1;
      }
      std::snprintf(buf, sizeof buf, c ? " %d %d" : "%d %d", r, sign);
      s += buf;
    }
    s += "\n";
  }
  s += "end\n";
  if (!out) return (int)s.size();
  if (s.size() + 1 > cap) return -4;
  s.copy(out, s.size());
  out[s.size()] = 0;
  return (int)s.size();
}

int chz_tc_verify(int n, int k, const long long *entries, const int *rotations, int t) {
  Mat W;
  if (!load_unitary(n, k, entries, W) || t < 0 || (t && !rotations)) return -3;
  const int P = 1 << (2 * n);
  // U = X_1 X_2 ... X_t C: peel X_1 first, each X = R(P_p) (inv 0) or R(P_p)^-1 (inv 1), by its inverse
  for (int i = 0; i < t; i++) {
    const int p = rotations[2 * i], inv = rotations[2 * i + 1];
    if (p < 1 || p >= P || (inv != 0 && inv != 1)) return -3;
    W = matmul(rotation(n, p, inv == 0), W);
  }
  // a unitary is a Clifford exactly when its transfer matrix is a signed permutation (every Pauli goes to +-1
  // times a Pauli): exponent 0, entries 0 or +-1, one nonzero per row and per column
  int e = 0;
  std::vector<std::pair<long long, long long>> xy;
  if (!transfer_matrix(n, W, e, xy)) return -2;
  if (e != 0) return 0;
  std::vector<int> rows(P, 0), cols(P, 0);
  for (int a = 0; a < P; a++)
    for (int b = 0; b < P; b++) {
      const auto v = xy[(size_t)a * P + b];
      if (v.second != 0 || v.first < -1 || v.first > 1) return 0;
      if (v.first) { rows[a]++; cols[b]++; }
    }
  for (int i = 0; i < P; i++)
    if (rows[i] != 1 || cols[i] != 1) return 0;
  return 1;
}

}  // extern "C"
