// Per-child device routines of the gate-count search.  A lane group of GS = 4^n lanes holds one residual
// matrix, lane le = r * N + c owning entry (r, c); for n = 3 a group is a whole wave64.
#pragma once
#include "dev.h"

namespace chz {

// For n = 3 a lane group is the whole wave, so values shared by a group are wave-uniform: reading them through
// readfirstlane keeps them in scalar registers (fewer vector registers live across the T bound).  Smaller groups
// share a wave with other groups and keep such values per lane.
template <int NQ> __device__ __forceinline__ int wuni(int v) {
  if constexpr (NQ == 3) return __builtin_amdgcn_readfirstlane(v);
  else return v;
}
template <int NQ> __device__ __forceinline__ long long wuni64(long long v) {
  if constexpr (NQ == 3)
    return (long long)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((int)v));
  else return v;
}

enum : uint8_t { ST_DEAD = 0, ST_ALIVE = 1, ST_FOUND = 2, ST_W2 = 3, ST_W3 = 4, ST_OVF = 6 };
constexpr uint32_t INFO_NOBOUND = 1u << 31;    // spec without bounds: every child is kept

// x * w^p for a coefficient vector of int32 values (entries below 2^28 in magnitude: no overflow)
__device__ __forceinline__ int4 zrot(int4 v, int p) {
  const int q = p & 3;
  const int nb = -v.y, nc = -v.z, nd = -v.w;
  // w (a, b, c, d) = (-d, a, b, c)
  int r0 = q == 0 ? v.x : q == 1 ? nd : q == 2 ? nc : nb;
  int r1 = q == 0 ? v.y : q == 1 ? v.x : q == 2 ? nd : nc;
  int r2 = q == 0 ? v.z : q == 1 ? v.y : q == 2 ? v.x : nd;
  int r3 = q == 0 ? v.w : q == 1 ? v.z : q == 2 ? v.y : v.x;
  const int m = (p & 4) ? -1 : 0;   // w^4 = -1
  return make_int4((r0 ^ m) - m, (r1 ^ m) - m, (r2 ^ m) - m, (r3 ^ m) - m);
}
__device__ __forceinline__ uint32_t uabs(int v) { return v < 0 ? 0u - (uint32_t)v : (uint32_t)v; }
__device__ __forceinline__ uint32_t zmag(int4 v) { return uabs(v.x) | uabs(v.y) | uabs(v.z) | uabs(v.w); }

// Child residual entry of lane le: (S g^-1)[r][c] = sum_m S[r][x_m] ginv[m][lc], then the group lowers the
// sqrt2 exponent while every entry is divisible by sqrt2 (src/cheese.cu:146-174, :773-817).  Overflow is
// judged before normalisation, where the reference narrows to int32.  Gates whose entries are all units
// +-w^p (every standard gate) rotate coefficients in int32 while the parent's entries stay below 2^28;
// otherwise the update runs in int64 with the general Z[w] product.  Returns false on overflow.
template <int NQ>
__device__ __forceinline__ bool make_child(const int4 *__restrict__ par, int pk, const DevGate &g, int le, int4 &s,
                                           int &k) {
  constexpr int N = 1 << NQ, GS = N * N;
  const int r = le >> NQ, c = le & (N - 1);
  const int nq = g.nq, L = g.L, mask = g.mask;
  const int q0 = g.q[0], q1 = g.q[1], q2 = g.q[2];
  int lc = (c >> q0) & 1;
  if (nq > 1) lc |= ((c >> q1) & 1) << 1;
  if (nq > 2) lc |= ((c >> q2) & 1) << 2;
  const int base = c & ~mask;
  bool of = false;
  const bool small = gall<GS>(zmag(par[r * N + c]) < (1u << 28));
  if (g.unit && small) {
    int4 acc = make_int4(0, 0, 0, 0);
    for (int m = 0; m < L; m++) {
      const int p = g.ph[m * L + lc];
      if (p < 0) continue;
      int x = base | ((m & 1) << q0);
      if (nq > 1) x |= ((m >> 1) & 1) << q1;
      if (nq > 2) x |= ((m >> 2) & 1) << q2;
      const int4 t = zrot(par[r * N + x], p);
      acc.x += t.x; acc.y += t.y; acc.z += t.z; acc.w += t.w;
    }
    s = acc;
  } else {
    long long a = 0, b = 0, cc = 0, d = 0;
    for (int m = 0; m < L; m++) {
      const int4 gm = g.inv[m * L + lc];
      if (zzero(gm)) continue;
      int x = base | ((m & 1) << q0);
      if (nq > 1) x |= ((m >> 1) & 1) << q1;
      if (nq > 2) x |= ((m >> 2) & 1) << q2;
      const int4 v = par[r * N + x];
      a += (long long)v.x * gm.x - (long long)v.y * gm.w - (long long)v.z * gm.z - (long long)v.w * gm.y;
      b += (long long)v.x * gm.y + (long long)v.y * gm.x - (long long)v.z * gm.w - (long long)v.w * gm.z;
      cc += (long long)v.x * gm.z + (long long)v.y * gm.y + (long long)v.z * gm.x - (long long)v.w * gm.w;
      d += (long long)v.x * gm.w + (long long)v.y * gm.z + (long long)v.z * gm.y + (long long)v.w * gm.x;
    }
    of = a != (int)a || b != (int)b || cc != (int)cc || d != (int)d;
    s = make_int4((int)a, (int)b, (int)cc, (int)d);
  }
  k = pk + g.ik;
  if (gany<GS>(of)) return false;
  while (k > 0 && gall<GS>(zhalf(s))) { s = zdown(s); k--; }
  return true;
}

// 2 if the matrix is monomial with a column permutation that is not affine over GF(2) (src/cheese.cu:321-338);
// b = group ballot of the nonzero entries
template <int NQ> __device__ __forceinline__ int h_struct(uint64_t b) {
  constexpr int N = 1 << NQ;
  uint64_t col0 = 0;
#pragma unroll
  for (int r = 0; r < N; r++) col0 |= 1ull << (r * N);
  int pi[N];
#pragma unroll
  for (int c = 0; c < N; c++) {
    const uint64_t cb = (b >> c) & col0;
    if (__popcll(cb) != 1) return 0;
    pi[c] = (__ffsll((long long)cb) - 1) / N;
  }
#pragma unroll
  for (int j = 0; j < N; j++) {
    int x = pi[0];
#pragma unroll
    for (int i = 0; i < NQ; i++) if ((j >> i) & 1) x ^= pi[1 << i] ^ pi[0];
    if (pi[j] != x) return 2;
  }
  return 0;
}

// entries (r, c) with bit q of r and c different, as a group-local lane mask
template <int NQ> __host__ __device__ constexpr uint64_t offdiag_mask(int q) {
  constexpr int N = 1 << NQ, GS = N * N;
  uint64_t m = 0;
  for (int le = 0; le < GS; le++)
    if ((((le / N) ^ (le % N)) >> q) & 1) m |= 1ull << le;
  return m;
}

// number of qubits the matrix acts on non-trivially (src/cheese.cu:305-319); b = group ballot of the nonzero
// entries.  A nonzero entry off the qubit's diagonal blocks settles that qubit without comparing blocks.
template <int NQ> __device__ __forceinline__ int support(int4 s, int le, uint64_t b) {
  constexpr int N = 1 << NQ, GS = N * N;
  const int r = le >> NQ, c = le & (N - 1);
  int cnt = 0;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    if (b & offdiag_mask<NQ>(q)) { cnt++; continue; }
    const int bit = 1 << q;
    const int4 o = gxor<GS>(s, bit * N + bit);
    const bool viol = !((r | c) & bit) && !zeq(s, o);
    cnt += gany<GS>(viol);
  }
  return cnt;
}

__host__ __device__ inline int spread_bits(int v, int mask, int n) {   // bit j of v -> j-th set bit of mask
  int out = 0, j = 0;
  for (int q = 0; q < n; q++)
    if ((mask >> q) & 1) { out |= ((v >> j) & 1) << q; j++; }
  return out;
}

// rank over F_p of the operator-Schmidt matrix across cut A (|A| = sa), R = 4^sa rows, counting up to `stop`
// (src/cheese.cu:270-303).  Lane le holds matrix entry (le / C, le % C).  Full pivoting: any nonzero entry
// of a row not yet used becomes the pivot and its column is cleared from the other unused rows by
// fraction-free updates (each a nonzero multiple of an F_p row combination), so after r steps the rank
// is r plus that of the unused rows; the rank over F_p does not depend on the pivot order.
// Lane le of the matrix across cut A (|A| = sa) holds entry (le >> lgC, le & (C - 1)), C = 2^lgC = 4^(n - sa):
// row (output bits, input bits) on A, column the same on the other qubits.  The residual entry it reads:
template <int NQ> __host__ __device__ inline int cut_source(int le, int A, int sa) {
  constexpr int N = 1 << NQ;
  const int lgC = 2 * (NQ - sa);
  const int mrow = le >> lgC, mcol = le & ((1 << lgC) - 1);
  const int nb = NQ - sa;
  const int ia = mrow & ((1 << sa) - 1), oa = mrow >> sa, ib = mcol & ((1 << nb) - 1), ob = mcol >> nb;
  const int notA = ((1 << NQ) - 1) & ~A;
  const int o = spread_bits(oa, A, NQ) | spread_bits(ob, notA, NQ);
  const int i = spread_bits(ia, A, NQ) | spread_bits(ib, notA, NQ);
  return o * N + i;
}

// src: cut_source of this lane, staged once per kernel; rows and columns are split by shifts (C is a power of 2)
template <int NQ> __device__ __forceinline__ int cut_rank(uint32_t fp, int le, int src, int sa, int stop) {
  constexpr int N = 1 << NQ, GS = N * N;
  (void)N;
  const int lgC = 2 * (NQ - sa), cm = (1 << lgC) - 1;
  const int mrow = le >> lgC, mcol = le & cm;
  uint32_t v = __shfl(fp, src, GS);
  int used = 0, rank = 0;                        // rows already pivots
  while (rank < stop) {
    const bool mine = !((used >> mrow) & 1);
    const uint64_t bb = gballot<GS>(mine && v != 0);
    if (!bb) break;
    if (rank + 1 == stop) return stop;          // the rank reaches the cap: no need to eliminate this pivot
    const int p = __ffsll((long long)bb) - 1;
    const int prow = p >> lgC, pcol = p & cm;
    const uint32_t pv = __shfl(v, p, GS);
    const uint32_t fr = __shfl(v, (mrow << lgC) + pcol, GS);
    const uint32_t pr = __shfl(v, (prow << lgC) + mcol, GS);
    used |= 1 << prow;
    if (mine && mrow != prow) v = fp_sub(mont_mul(v, pv), mont_mul(fr, pr));
    rank++;
  }
  return rank;
}

// Bound word carried by every frontier node: bits 0-7 its raw T bound (BW_TUNK: not known), bits 8 + 3 i the
// need of cut i (ceil log2 of its clipped rank, at most 2 for n <= 3), bit 30 set when one of its children is
// the goal, bit 31 set when the cut needs are known.  Roots carry BW_TUNK.  A child copies what its gate cannot
// change and computes the rest.
constexpr uint32_t BW_TUNK = 0xffu, BW_CUTS = 1u << 31, BW_GOAL = 1u << 30, BW_NEEDS = 0x1fffff00u;
constexpr int BW_MAXCUTS = 7;

// CX/CZ lower bound from the cut ranks and the ILP table (src/cheese.cu:585-597).  Cuts in `keepc` take the
// need packed in `needs` (the parent's: the gate acts on one side of them, so the rank is unchanged); the
// others are ranked here.  On return `needs` holds this matrix's packed needs.
// The cx table is read from the expansion kernel's LDS copy `cxl` when it fits (sp.cxlds): the global table sits
// behind a vector load on every child's dependency chain.
__device__ __forceinline__ int cx_at(const KSpec &sp, const uint8_t *cxl, int idx) {
  return sp.cxlds ? (int)cxl[idx] : sp.cxtable[idx];
}

template <int NQ> __device__ __forceinline__ int cx_bound(int4 s, int le, const KSpec &sp, const uint8_t *cxl,
                                                          const uint8_t *csrc, uint32_t keepc, uint32_t &needs) {
  constexpr int GS = 1 << (2 * NQ);
  uint32_t fp = 0;
  bool have_fp = false;
  int idx = 0, mult = 1, mx = 0;
  uint32_t out = 0;
  for (int i = 0; i < sp.ncuts; i++) {
    const int rad = sp.rad[i], sa = sp.cutsize[i];
    int need;
    if ((keepc >> i) & 1) need = (int)((needs >> (3 * i)) & 7u);
    else {
      if (!have_fp) { fp = fp_of(s, sp.wm); have_fp = true; }
      const int R = 1 << (2 * sa);
      int stop = R;
      if (rad >= 2 && rad - 2 < 30 && (1 << (rad - 2)) + 1 < stop) stop = (1 << (rad - 2)) + 1;
      const int src = i < CSRC_MAX ? (int)csrc[i * GS + le] : cut_source<NQ>(le, sp.cutmask[i], sa);
      const int rk = cut_rank<NQ>(fp, le, src, sa, stop);
      need = 0;
      while ((1 << need) < rk) need++;
      if (need > rad - 1) need = rad - 1;
    }
    if (i < BW_MAXCUTS) out |= (uint32_t)need << (3 * i);
    idx += need * mult;
    mult *= rad;
    if (need > mx) mx = need;
  }
  needs = out;
  return sp.have_table ? cx_at(sp, cxl, idx) : mx;
}

// Lower bound on a child's cut bound from its parent's packed needs: each need drops by at most the gate's
// per-cut delta (operator-Schmidt ranks multiply), and the table is non-decreasing in every need.
__device__ __forceinline__ int cx_lower(const KSpec &sp, const uint8_t *cxl, uint32_t pneeds, uint32_t cdel) {
  int idx = 0, mult = 1, mx = 0;
  for (int i = 0; i < sp.ncuts; i++) {
    const int dl = (int)((cdel >> (3 * i)) & 7u);
    int nd = (int)((pneeds >> (3 * i)) & 7u) - dl;
    nd = dl == 7 || nd < 0 ? 0 : nd;
    idx += nd * mult;
    mult *= sp.rad[i];
    if (nd > mx) mx = nd;
  }
  return sp.have_table ? cx_at(sp, cxl, idx) : mx;
}

// T lower bound: max over Pauli pairs of 2(n+k) - v(X_ab), all X_ab by one 2-D Walsh-Hadamard transform
// per flip pair f = (fa, fb) of Q_f[e] = S[e ^ f] conj(S[e]), arithmetic in Z[w] mod 2^bits(U)
// (src/cheese.cu:397-449).  The minimum valuation over all transforms is read from two OR accumulators:
// O = OR of all coefficients, D = OR of (a ^ c) | (b ^ d); min v = 2 m + [bit m of D clear], m = ctz(O).
template <int NQ, typename U> __device__ __forceinline__ int t_bound(int4 s, int k, int le) {
  constexpr int N = 1 << NQ, GS = N * N;
  const U ya = (U)(long long)s.x, yb = (U)0 - (U)(long long)s.w, yc = (U)0 - (U)(long long)s.z,
          yd = (U)0 - (U)(long long)s.y;
  U O = 0, D = 0;
  for (int f = 0; f < GS; f++) {
    const int4 xs = gxor<GS>(s, f);
    const U xa = (U)(long long)xs.x, xb = (U)(long long)xs.y, xc = (U)(long long)xs.z, xd = (U)(long long)xs.w;
    U qa = U(xa * ya - xb * yd - xc * yc - xd * yb);
    U qb = U(xa * yb + xb * ya - xc * yd - xd * yc);
    U qc = U(xa * yc + xb * yb + xc * ya - xd * yd);
    U qd = U(xa * yd + xb * yc + xc * yb + xd * ya);
#pragma unroll
    for (int h = 1; h < GS; h <<= 1) {
      const U oa = __shfl_xor(qa, h, GS), ob = __shfl_xor(qb, h, GS), oc = __shfl_xor(qc, h, GS),
              od = __shfl_xor(qd, h, GS);
      if (le & h) { qa = U(oa - qa); qb = U(ob - qb); qc = U(oc - qc); qd = U(od - qd); }
      else { qa = U(qa + oa); qb = U(qb + ob); qc = U(qc + oc); qd = U(qd + od); }
    }
    O |= qa | qb | qc | qd;
    D |= (qa ^ qc) | (qb ^ qd);
  }
  O = gor<GS>(O);
  D = gor<GS>(D);
  if (!O) return 0;
  const int m = __ffsll((unsigned long long)O) - 1;
  const int v = 2 * m + (((D >> m) & 1) ? 0 : 1);
  const int need = 2 * (NQ + k) - v;
  return need > 0 ? need : 0;
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 pk(int lo, int hi) { u16x2 r; r.x = (unsigned short)lo; r.y = (unsigned short)hi; return r; }
__device__ __forceinline__ uint32_t bits(u16x2 v) { return (uint32_t)v.x | ((uint32_t)v.y << 16); }

// Per-entry operands of the packed transform, staged in LDS by the lane owning entry e, one array per field
// (a group's entries are read at a stride of 8 bytes, free of bank conflicts): x[e] = (a, b | c, d) of S[e] mod
// 2^16, and p[pass * TS_STRIDE + e] the four packed multipliers of conj(S[e]) = (a, -d, -c, -b) for a pass (one
// 128-bit read per entry and pass).  With
// Q = S[x] conj(S[e]) = (Qa, Qb, Qc, Qd) the transform needs only Qa, Qc and Qb -+ Qd (below), in two passes:
//   (Qa | Qc)           = xa (a | -c)    + xb (b | -d)    + xc (c | a)     + xd (d | b)
//   (Qb - Qd | Qb + Qd) = xa (b-d | -b-d) + xb (a+c | a-c) + xc (b+d | b-d) + xd (c-a | c+a)
constexpr int TS_STRIDE = 256;      // entries per field: one per thread of a block
struct TStage {
  uint2 *x;
  uint4 *p;
};

template <int NQ> __device__ __forceinline__ void t_stage(TStage st, int4 s, int le) {
  const int a = s.x, b = s.y, c = s.z, d = s.w;
  st.x[le] = make_uint2(bits(pk(a, b)), bits(pk(c, d)));
  st.p[le] = make_uint4(bits(pk(a, -c)), bits(pk(b, -d)), bits(pk(c, a)), bits(pk(d, b)));
  st.p[TS_STRIDE + le] = make_uint4(bits(pk(b - d, -b - d)), bits(pk(a + c, a - c)), bits(pk(b + d, b - d)),
                                    bits(pk(c - a, c + a)));
}

__device__ __forceinline__ u16x2 as16(uint32_t v) { u16x2 r; r.x = (unsigned short)v; r.y = (unsigned short)(v >> 16); return r; }

__device__ __forceinline__ int ctz16(uint32_t v) { return v & 0xffffu ? __ffs((int)(v & 0xffffu)) - 1 : 16; }

// T bound with arithmetic mod 2^16 (exact when n + k <= 16; a value that is 0 mod 2^16 has valuation >= 32 >=
// 2(n+k) and cannot raise the bound).  Q_f[e] = S[e ^ f] conj(S[e]) satisfies Q_f[e ^ f] = conj(Q_f[e]), so for
// f != 0, with j the top bit of f and R = {e : e_j = 0}, the 2-D transform is
//   W_f[z] = A[z] +- conj(A[z]),   A[z] = sum_{e in R} (-1)^{e.z} Q_f[e]   (a transform over the other bits),
// and A + conj A = 2a + (b - d) sqrt2, A - conj A = i (2c + (b + d) sqrt2) for A = (a, b, c, d).  An element
// x + y sqrt2 has valuation min(2 v2(x), 2 v2(y) + 1), so the smallest valuation v over all outputs is the
// minimum of three parts, each read from an OR accumulator:
//   f = 0 (Q_0 = |S|^2 = r + y sqrt2, real; transformed by the whole group, one lane per entry): 2 v2(r),
//       1 + 2 v2(y)
//   pass 1, (b - d | b + d) of every A:  1 + 2 v2, at least 1
//   pass 0, (a | c) of every A:          2 + 2 v2, at least 2
// and the bound is 2(n + k) - v.  Each part only lowers v, so 2(n + k) minus the minimum over the parts computed
// so far is a lower bound on the T bound, exact once no remaining part can go below it.  A pass is the costly
// part (E/2 Z[w] products and a transform per lane); the expansion runs them only for children the lower bound
// leaves undecided, the odd pass first (it sets the minimum far more often).

// The staged T bound: the f = 0 part, then pass 1, then pass 0, each only while the lower bound
// max(Tlo, 2(n + k) - v) stays at most tmax (the largest T bound at which the child is alive) and a remaining part
// could still lower v.  Tlo is a lower bound the caller already has (the parent's bracket, a sibling's value).
// Returns the bound reached; exact tells whether it is the T bound (always so when it is at most tmax).  npass
// counts the passes run.  One function with the pass loop inside keeps the pass's LDS reads interleaved with its
// products; split into separate helpers the compiler issued all 64 reads of a pass at once and spilled.
template <int NQ> __device__ __forceinline__ int t_bound16s(TStage st, int4 s, int k, int le, int Tlo, int tmax,
                                                             uint32_t &npass, bool &exact) {
  constexpr int E = 1 << (2 * NQ), H = E / 2, GS = E;
  const int top = 2 * (NQ + k);
  // f = 0: |S[e]|^2 = r + y sqrt2 with r = a^2 + b^2 + c^2 + d^2, y = ab + bc + cd - da
  uint32_t r = (uint32_t)(s.x * s.x + s.y * s.y + s.z * s.z + s.w * s.w);
  uint32_t y = (uint32_t)(s.x * s.y + s.y * s.z + s.z * s.w - s.w * s.x);
#pragma unroll
  for (int h = 1; h < GS; h <<= 1) {
    const uint32_t ro = __shfl_xor(r, h, GS), yo = __shfl_xor(y, h, GS);
    if (le & h) { r = ro - r; y = yo - y; } else { r = r + ro; y = y + yo; }
  }
  // group ORs are uniform over the group: read through readfirstlane (n = 3) they stay in scalar registers
  const uint32_t accG = (uint32_t)wuni<NQ>((int)gor<GS>(r & 0xffffu));
  const uint32_t accY = (uint32_t)wuni<NQ>((int)gor<GS>(y & 0xffffu));
  int v = 2 * ctz16(accG);
  const int vy = 1 + 2 * ctz16(accY);
  v = vy < v ? vy : v;
  int T = top - v > Tlo ? top - v : Tlo;
  exact = v <= 1;
  if (exact || T > tmax) return T;
  __builtin_amdgcn_wave_barrier();
  t_stage<NQ>(st, s, le);
  __builtin_amdgcn_wave_barrier();
  // lane f of the group handles flip pair f (lane 0 repeats f = 1; OR-accumulation is idempotent)
  int f = le == 0 ? 1 : le;
  // keep the per-lane LDS address arithmetic inside the child loop: hoisted, its 64 addresses would
  // halve the kernel's occupancy
  asm volatile("" : "+v"(f));
  const int j = 31 - __clz(f);
  const int lowm = (1 << j) - 1;
#pragma unroll 1
  for (int pass = 1; pass >= 0; pass--) {
    const uint4 *pm = st.p + pass * TS_STRIDE;
    // the addresses below depend on a value the compiler cannot see through, so they are formed per pass
    // instead of being hoisted out of the pass loop (64 more live registers)
    int fp = f, lm = lowm;
    asm volatile("" : "+v"(fp), "+v"(lm));
    u16x2 q[H];
#pragma unroll
    for (int i = 0; i < H; i++) {
      const int e = ((i & ~lm) << 1) | (i & lm);         // i with a 0 inserted at bit j
      const uint2 xv = st.x[e ^ fp];
      const u16x2 x01 = as16(xv.x), x23 = as16(xv.y);
      const uint4 m = pm[e];
      q[i] = x01.xx * as16(m.x) + x01.yy * as16(m.y) + x23.xx * as16(m.z) + x23.yy * as16(m.w);
    }
#pragma unroll
    for (int h = 1; h < H; h <<= 1)
#pragma unroll
      for (int i = 0; i < H; i++)
        if (!(i & h)) {
          const u16x2 u = q[i], w = q[i | h];
          q[i] = u + w;
          q[i | h] = u - w;
        }
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < H; i++) o |= bits(q[i]);
    const uint32_t acc = (uint32_t)wuni<NQ>((int)gor<GS>((o & 0xffffu) | (o >> 16)));
    npass++;
    const int vp = (pass ? 1 : 2) + 2 * ctz16(acc);    // pass 1: 1 + 2 v2 of (b - d | b + d); pass 0: 2 + 2 v2 of (a | c)
    v = vp < v ? vp : v;
    T = top - v > Tlo ? top - v : Tlo;
    exact = pass == 0 || v <= 2;
    if (exact || T > tmax) break;
  }
  __builtin_amdgcn_wave_barrier();
  return T;
}

template <int NQ> __device__ __forceinline__ int t_bound_any(int4 s, int k, int le) {
  return 2 * (NQ + k) <= 64 ? t_bound<NQ, uint32_t>(s, k, le) : t_bound<NQ, unsigned long long>(s, k, le);
}

// 32-bit hash of one entry's four coefficients.  A multiply moves information only towards the high bits, so
// each round folds the high half back down before the next coefficient enters; without that, entries with small
// coefficients collide far more often than a random function would.
__device__ __forceinline__ uint32_t entry_hash32(int4 v) {
  uint32_t h = (uint32_t)v.x * 0x9e3779b1u + (uint32_t)v.y;
  h ^= h >> 15;
  h = h * 0x85ebca77u + (uint32_t)v.z;
  h ^= h >> 13;
  h = h * 0xc2b2ae3du + (uint32_t)v.w;
  h ^= h >> 16;
  h *= 0x27d4eb2fu;
  return h ^ (h >> 15);
}
// the 64-bit summand of an entry hash at position pos (1..64): two independent one-multiply mixes of the pair.
// 32-bit multiplies run at a quarter of the VALU rate and 64-bit ones take four of them, so the summand costs
// two where a 64-bit mix costs eight; the sum over the entries still differs between two matrices unless their
// differing entries' hashes collide (2^-32 per entry) or the 64-bit sums do.
__device__ __forceinline__ uint64_t slot64(uint32_t h, uint32_t pos) {
  const uint32_t x = h ^ __umul24(pos, 0x9e3779u);
  uint32_t lo = x ^ (x >> 16);
  lo *= 0x7feb352du;
  lo ^= lo >> 15;
  uint32_t hi = x ^ 0x5bd1e995u;
  hi ^= hi >> 15;
  hi *= 0x2c1b3c6du;
  hi ^= hi >> 12;
  return ((uint64_t)hi << 32) | lo;
}

// class key: the smallest over the symmetry group of an entry-sum hash of sigma(S); equal classes give
// equal keys, and equality of keys is always confirmed exactly (src/cheese.cu:224-241)
template <int NQ> __device__ __forceinline__ uint64_t class_key(int4 s, int k, int le, const KSpec &sp) {
  constexpr int N = 1 << NQ, GS = N * N;
  const int r = le >> NQ, c = le & (N - 1);
  const uint32_t hv = entry_hash32(s), hc = entry_hash32(zconj(s));
  uint64_t best = ~0ull;
  for (int y = 0; y < sp.nsym; y++) {
    const uint8_t *pm = sp.symperm + y * N;
    const uint32_t pos = (uint32_t)(pm[r] * N + pm[c]) + 1;
    const uint64_t h = slot64(sp.symconj[y] ? hc : hv, pos);
    const uint64_t key = mix64(gsum<GS>(h) ^ ((uint64_t)(k + 1) * 0xc2b2ae3d27d4eb4full));
    best = key < best ? key : best;
  }
  return best;
}

// ---- wave64 exchanges without the LDS crossbar (gfx950): permlane32/16 swaps and DPP row controls ----
__device__ __forceinline__ uint64_t u64_of(uint32_t lo, uint32_t hi) { return (uint64_t)lo | ((uint64_t)hi << 32); }

// lanes 0-31 receive a + a[lane ^ 32], lanes 32-63 receive b + b[lane ^ 32]: v_permlane32_swap exchanges the
// upper half of its first operand with the lower half of its second
__device__ __forceinline__ uint64_t rs32(uint64_t a, uint64_t b) {
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)a, (uint32_t)b, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(a >> 32), (uint32_t)(b >> 32), false, false);
  return u64_of(lo[0], hi[0]) + u64_of(lo[1], hi[1]);
}
// lanes with bit 4 clear receive a + a[lane ^ 16], the others b + b[lane ^ 16]: v_permlane16_swap exchanges
// the odd 16-lane rows of its first operand with the even rows of its second
__device__ __forceinline__ uint64_t rs16(uint64_t a, uint64_t b) {
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)a, (uint32_t)b, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(a >> 32), (uint32_t)(b >> 32), false, false);
  return u64_of(lo[0], hi[0]) + u64_of(lo[1], hi[1]);
}
__device__ __forceinline__ uint64_t partner32(uint64_t x, int lane) {       // x[lane ^ 32]
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)x, (uint32_t)x, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(x >> 32), (uint32_t)(x >> 32), false, false);
  return (lane & 32) ? u64_of(lo[0], hi[0]) : u64_of(lo[1], hi[1]);
}
__device__ __forceinline__ uint64_t partner16(uint64_t x, int lane) {       // x[lane ^ 16]
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)x, (uint32_t)x, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(x >> 32), (uint32_t)(x >> 32), false, false);
  return (lane & 16) ? u64_of(lo[0], hi[0]) : u64_of(lo[1], hi[1]);
}
template <int CTRL> __device__ __forceinline__ uint64_t dpp64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xf, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xf, 0xf, false);
  return u64_of(lo, hi);
}

// class_key for n = 3 (one matrix per wave64, whole wave active), at most 16 symmetries: the 16 per-lane
// hashes are reduce-scattered over the wave (lane ^ 32, ^ 16 by permlane swaps; ^ 15, ^ 7 by DPP mirrors; then
// sums over each quad), so lane L ends with the wave sum of slot (L >> 2) & 15; one mix per lane and a min over
// lane bits 2-5 give the same key as class_key.  spos[y * 64 + le]: position of entry le under symmetry y (any
// entry index for y >= nsym: those slots are hashed too, and their lanes are left out of the final min);
// cjm bit y: symmetry y conjugates.  The conjugation bits are tested on a vector copy of cjm: tested on the
// scalar they become 16 loop-invariant lane masks that the compiler hoists out of the child loop, where they
// spill.
__device__ __forceinline__ uint64_t class_key3(int4 s, int k, int lane, const uint8_t *spos, uint32_t cjm, int nsym) {
  const uint32_t hv = entry_hash32(s), hc = entry_hash32(zconj(s));
  uint32_t cjv = cjm;
  asm volatile("" : "+v"(cjv));
  auto slot = [&](int y) -> uint64_t { return slot64(((cjv >> y) & 1) ? hc : hv, (uint32_t)spos[y * 64 + lane] + 1); };
  uint64_t h[8];      // slots y and y + 8 are combined as soon as they are hashed: 8 live values, not 16
#pragma unroll
  for (int j = 0; j < 8; j++) h[j] = rs32(slot(j), slot(j + 8));
#pragma unroll
  for (int j = 0; j < 4; j++) h[j] = rs16(h[j], h[j + 4]);
  const bool b3 = lane & 8, b2 = lane & 4;
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const uint64_t send = b3 ? h[j] : h[j + 2], keep = b3 ? h[j + 2] : h[j];
    h[j] = keep + dpp64<DPP_MIRROR>(send);
  }
  uint64_t x;
  {
    const uint64_t send = b2 ? h[0] : h[1], keep = b2 ? h[1] : h[0];
    x = keep + dpp64<DPP_HALF_MIRROR>(send);
  }
  x += dpp64<DPP_QXOR1>(x);
  x += dpp64<DPP_QXOR2>(x);
  const int y = (lane >> 2) & 15;
  uint64_t key = y < nsym ? mix64(x ^ ((uint64_t)(k + 1) * 0xc2b2ae3d27d4eb4full)) : ~0ull;
  uint64_t o = dpp64<DPP_HALF_MIRROR>(key);
  key = o < key ? o : key;
  o = dpp64<DPP_MIRROR>(key);
  key = o < key ? o : key;
  o = partner16(key, lane);
  key = o < key ? o : key;
  o = partner32(key, lane);
  return o < key ? o : key;
}

// exact: B = sigma(A) for some group element (src/cheese.cu:242-255)
// the same test with the group's entry maps staged in LDS: spos[y * E + e] is the entry of b that lands on entry e
// under group element y, bit y of cjm its conjugation flag
template <int NQ> __device__ __forceinline__ bool same_class_lds(int4 a, int ka, int4 b, int kb, int le,
                                                                const uint8_t *spos, uint32_t cjm, int nsym) {
  constexpr int N = 1 << NQ, GS = N * N;
  if (ka != kb) return false;
  const int4 ac = zconj(a);
  for (int y = 0; y < nsym; y++) {
    const int4 bv = gshfl<GS>(b, spos[y * GS + le]);
    const int4 av = (cjm >> y) & 1 ? ac : a;
    if (gall<GS>(zeq(av, bv))) return true;
  }
  return false;
}

// the same test against a residual held in LDS (rep: E entries), read through the staged maps
template <int NQ> __device__ __forceinline__ bool same_class_rep(int4 a, int ka, const int4 *rep, int kb, int le,
                                                                const uint8_t *spos, uint32_t cjm, int nsym) {
  constexpr int GS = (1 << NQ) * (1 << NQ);
  if (ka != kb) return false;
  const int4 ac = zconj(a);
  for (int y = 0; y < nsym; y++) {
    const int4 bv = rep[spos[y * GS + le]];
    const int4 av = (cjm >> y) & 1 ? ac : a;
    if (gall<GS>(zeq(av, bv))) return true;
  }
  return false;
}

template <int NQ> __device__ __forceinline__ bool same_class(int4 a, int ka, int4 b, int kb, int le, const KSpec &sp) {
  constexpr int N = 1 << NQ, GS = N * N;
  if (ka != kb) return false;
  const int r = le >> NQ, c = le & (N - 1);
  for (int y = 0; y < sp.nsym; y++) {
    const uint8_t *pm = sp.symperm + y * N;
    const int4 bv = gshfl<GS>(b, pm[r] * N + pm[c]);
    const int4 av = sp.symconj[y] ? zconj(a) : a;
    if (gall<GS>(zeq(av, bv))) return true;
  }
  return false;
}

}  // namespace chz
