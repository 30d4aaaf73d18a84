// Optimal T-count on one MI355X (reference behaviour: src/tcount.cu): a coset table grown level by level from
// the identity into an HBM open-addressing table, and a batched search that peels rotations off many targets at
// once and looks every residual up in that table.
//
// A node is a right coset S C of a channel matrix S (P x P, P = 4^n, entries (x + y sqrt2) / sqrt2^e); its
// children peel one pi/8 rotation R(P_p), p = 1 .. P-1, off the left (src/tcount.cu:80-112).  Nodes are kept as
// (parent, rotation) pairs per level.  A level that is about to be expanded is materialised in HBM (P*P words per
// node, row-major, so that lane c reading column c of every row is one coalesced 256-byte access per row);
// otherwise a wave replays a node's path from the nearest materialised level.  Each lane owns one column:
// peeling acts on rows, so a lane's column never leaves its lane.
//
// Representation: every entry is held as its value in F_p, p = 2^31 - 1, where sqrt2 = 2^16 (2^32 = 2 mod p)
// and 1/sqrt2 = 2^15, so scaling by sqrt2 is a 31-bit rotation and a residue does not depend on the exponent
// chosen (no renormalisation).  This is exact, not a hash: a channel matrix is orthogonal and so is its Galois
// conjugate (sqrt2 -> -sqrt2), so at any common exponent e both x + y sqrt2 and x - y sqrt2 lie in
// [-sqrt2^e, sqrt2^e]; for e <= 28 that bounds the integer x + 2^16 y strictly inside (-p/2, p/2), where the
// residue determines it, and |x| < 2^15 then determines (x, y).  Two nodes whose exponents are <= 28 are
// therefore equal exactly when their residues are.  The host checks the exponent bound before a run.
//  * keys: a column hashes to (sum_r R[r] v_r)^2 (the square removes the column sign); the coset key is a mixed
//    sum over columns (removes the column order).  A child's column hash is a fixed linear form of its
//    parent's column (coefficients per rotation, precomputed), so all P-1 children of a parent come from one
//    parent column held in registers.
//  * equality (dedupe and table hits, src/tcount.cu:174-188): columns are sign-normalised (the residue of the
//    first nonzero entry made to stand for a positive value), sorted by a column hash and compared entry by
//    entry.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cheese_hip.h"
#include "host.h"
#include "primitives.h"

namespace tcn {

constexpr int MAXT = CHZ_TMAX;
constexpr int EMAX = 28;                    // largest exponent at which residues identify channel entries
constexpr uint32_t MP = 0x7fffffffu;        // 2^31 - 1
constexpr uint32_t HALF = MP / 2;           // residues above it stand for negative values
constexpr int DMAXREP = 32;                 // distinct cosets kept per run of equal key prefixes

// ---- F_p, p = 2^31 - 1
__host__ __device__ __forceinline__ uint32_t m31_red(uint64_t x) {
  x = (x & MP) + (x >> 31);
  x = (x & MP) + (x >> 31);
  return (uint32_t)(x >= MP ? x - MP : x);
}
__host__ __device__ __forceinline__ uint32_t m31_add(uint32_t a, uint32_t b) { const uint32_t s = a + b; return s >= MP ? s - MP : s; }
__host__ __device__ __forceinline__ uint32_t m31_sub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + MP - b; }
__host__ __device__ __forceinline__ uint32_t m31_rot(uint32_t v, int k) {      // v 2^k, v < p, 0 <= k < 31
  return k ? (((v << k) | (v >> (31 - k))) & MP) : v;
}
__host__ __device__ __forceinline__ uint32_t m31_of(int x) {
  if (x >= 0) return (uint32_t)x % MP;
  const uint32_t r = (uint32_t)(-(int64_t)x % MP);
  return r ? MP - r : 0;
}
// (x + y sqrt2) / sqrt2^e in F_p
__host__ __device__ __forceinline__ uint32_t m31_entry(int x, int y, int e) {
  return m31_rot(m31_add(m31_of(x), m31_rot(m31_of(y), 16)), (15 * e) % 31);
}
// sum_r c_r v_r for c_r, v_r < p: each product split as (c mod 2^16) v + (c >> 16) v 2^16, 64 terms in 64 bits
__device__ __forceinline__ uint32_t m31_fold(uint64_t lo, uint64_t hi) { return m31_add(m31_red(lo), m31_rot(m31_red(hi), 16)); }

__host__ __device__ __forceinline__ uint64_t tmix(uint64_t x) {
  x ^= x >> 32; x *= 0xd6e8feb86659fd93ull;
  x ^= x >> 32; x *= 0xd6e8feb86659fd93ull;
  return x ^ (x >> 32);
}
__device__ __forceinline__ uint64_t col_term(uint32_t g) { return tmix((uint64_t)g + 0x9e3779b97f4a7c15ull); }
__host__ __device__ __forceinline__ uint64_t coset_key(uint64_t s) { return tmix(s ^ 0x632be59bd9b4e019ull); }
// sort prefix: keys of different instances land in different runs
__host__ __device__ __forceinline__ uint32_t sort_prefix(uint64_t key, uint32_t inst) {
  return (uint32_t)((key ^ ((uint64_t)inst * 0x9e3779b97f4a7c15ull)) >> 32);
}

__device__ __forceinline__ uint64_t wsum64(uint64_t v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}
// rotation tables and hash coefficients on the device
struct Rot {
  const uint64_t *anti;      // (P-1): bit a set when row a anticommutes with P_p (its partner is row a ^ p)
  const uint64_t *plus;      // (P-1): bit a set when the sign s_a of row a is +1
  const uint32_t *coef;      // (P-1) x P: child column hash = sum_r coef[p][r] v_r over the parent's column
  const uint32_t *rv;        // P: column hash = sum_r rv[r] v_r
  const uint32_t *rw;        // P: second weights of the column hash used to pair columns
};

// nodes of every level: level 0 node i is root i; a node at level t >= 1 is (par, rot)
struct Forest {
  const uint32_t *par[MAXT + 1];
  const uint8_t *rot[MAXT + 1];
  const uint32_t *inst[MAXT + 1];   // instance of each node (null: one instance)
  const uint32_t *mat[MAXT + 1];    // materialised levels (node x P*P residues) or null; mat[0]: the roots
};

__device__ __forceinline__ uint32_t inst_of(const Forest &F, int t, uint32_t node) {
  return F.inst[t] ? F.inst[t][node] : 0u;
}

// one rotation peeled off M in place, lane c rewriting its column c: rows that commute with P_p keep their
// values, an anticommuting pair (a, a ^ p) becomes ((a - s_a r) / sqrt2, (r - s_r a) / sqrt2)
// (src/tcount.cu:80-99).  Pairs are enumerated as the rows a with the top bit of p clear.
template <int P>
__device__ __forceinline__ void fp_peel(uint32_t *M, const Rot &R, int p, int lane) {
  const uint64_t anti = R.anti[p - 1], plus = R.plus[p - 1];
  const int h = 31 - __clz(p), lowm = (1 << h) - 1;
  if (lane < P) {
#pragma unroll 4
    for (int j = 0; j < P / 2; j++) {
      const int a = ((j >> h) << (h + 1)) | (j & lowm), r = a ^ p;
      if ((anti >> a) & 1) {
        uint32_t *ma = M + a * P + lane, *mr = M + r * P + lane;
        const uint32_t va = *ma, vr = *mr;
        *ma = m31_rot(((plus >> a) & 1) ? m31_sub(va, vr) : m31_add(va, vr), 15);
        *mr = m31_rot(((plus >> r) & 1) ? m31_sub(vr, va) : m31_add(vr, va), 15);
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// P*P words from src (global) into M (LDS), loads issued in batches ahead of the stores
template <int P>
__device__ __forceinline__ void copy_in(const uint32_t *__restrict__ src, uint32_t *__restrict__ M, int lane) {
  constexpr int N = P * P / 64 > 0 ? P * P / 64 : 1, B = N < 16 ? N : 16;
  if (P * P < 64) {
    if (lane < P * P) M[lane] = src[lane];
  } else {
#pragma unroll
    for (int k0 = 0; k0 < N; k0 += B) {
      uint32_t t[B];
#pragma unroll
      for (int j = 0; j < B; j++) t[j] = src[(k0 + j) * 64 + lane];
#pragma unroll
      for (int j = 0; j < B; j++) M[(k0 + j) * 64 + lane] = t[j];
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// node (t, i) into M (LDS): copied from the nearest materialised level at or below t, then the remaining
// rotations of its path peeled off
template <int P>
__device__ void load_node(const Forest &F, const Rot &R, int t, uint32_t node, uint32_t *M, uint8_t *seq, int lane) {
  int l0 = t;
  uint32_t i = node;
  if (lane == 0)
    while (!F.mat[l0]) { seq[l0 - 1] = F.rot[l0][i]; i = F.par[l0][i]; l0--; }
  l0 = __shfl(l0, 0, 64);
  i = __shfl(i, 0, 64);
  __builtin_amdgcn_wave_barrier();
  copy_in<P>(F.mat[l0] + (size_t)i * (P * P), M, lane);
  for (int l = l0; l < t; l++) fp_peel<P>(M, R, __builtin_amdgcn_readfirstlane((int)seq[l]), lane);
}

// lane's column of M sign-normalised in place (the residue of its first nonzero entry made to stand for a positive
// value) and its sort key: a 58-bit hash of the normalised column above the 6-bit column index
template <int P>
__device__ uint64_t norm_col(uint32_t *M, const Rot &R, int lane) {
  if (lane >= P) return ~0ull;
  bool seen = false, neg = false;
  for (int r = 0; r < P; r++) {
    const uint32_t v = M[r * P + lane];
    if (!seen && v) { seen = true; neg = v > HALF; }
  }
  uint64_t lo1 = 0, hi1 = 0, lo2 = 0, hi2 = 0;
  for (int r = 0; r < P; r++) {
    uint32_t v = M[r * P + lane];
    if (neg && v) { v = MP - v; M[r * P + lane] = v; }
    const uint32_t k1 = R.rv[r], k2 = R.rw[r];
    lo1 += (uint64_t)(k1 & 0xffffu) * v;
    hi1 += (uint64_t)(k1 >> 16) * v;
    lo2 += (uint64_t)(k2 & 0xffffu) * v;
    hi2 += (uint64_t)(k2 >> 16) * v;
  }
  const uint64_t h = (uint64_t)m31_fold(lo1, hi1) | ((uint64_t)(m31_fold(lo2, hi2) & 0x7ffffffu) << 31);
  return (h << 6) | (uint32_t)lane;
}

__device__ __forceinline__ uint64_t bitonic64(uint64_t v, int lane) {   // ascending over the wave
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint64_t o = __shfl_xor(v, j, 64);
      const bool take_min = ((lane & j) == 0) == ((lane & k) == 0);
      v = take_min ? (o < v ? o : v) : (o > v ? o : v);
    }
  return v;
}

template <int P>
__device__ __forceinline__ bool col_eq(const uint32_t *A, int ca, const uint32_t *B, int cb) {
  for (int r = 0; r < P; r++)
    if (A[r * P + ca] != B[r * P + cb]) return false;
  return true;
}

// right-coset equality of two sign-normalised matrices (keys from norm_col): equal multisets of columns.
// Columns are paired after sorting by hash and compared entry by entry; if a hash tie paired different
// columns, every column's multiplicity is counted in both.
template <int P>
__device__ bool same_cols(const uint32_t *A, uint64_t ka, const uint32_t *B, uint64_t kb, int lane) {
  const uint64_t sa = bitonic64(ka, lane), sb = bitonic64(kb, lane);
  if (__ballot((sa >> 6) != (sb >> 6))) return false;
  const bool eq = lane >= P || col_eq<P>(A, (int)(sa & 63), B, (int)(sb & 63));
  if (!__ballot(!eq)) return true;
  bool ok = true;
  if (lane < P) {
    int na = 0, nb = 0;
    for (int c = 0; c < P; c++) { na += col_eq<P>(A, lane, A, c); nb += col_eq<P>(A, lane, B, c); }
    ok = na == nb;
  }
  return !__ballot(!ok);
}

// child (parent at level t, rotation p) of F, sign-normalised, into M; returns the lane's column key
template <int P>
__device__ uint64_t load_child(const Forest &F, const Rot &R, int t, uint32_t parent, int p, uint32_t *M,
                               uint8_t *seq, int lane) {
  load_node<P>(F, R, t, parent, M, seq, lane);
  if (p) fp_peel<P>(M, R, p, lane);
  const uint64_t k = norm_col<P>(M, R, lane);
  __builtin_amdgcn_wave_barrier();
  return k;
}

// ------------------------------------------------------------------------------------------------------
// kernels

struct RootKeyArgs {
  Rot R;
  const uint32_t *root_fp;
  long long n;
  uint64_t *keys;
};

// coset keys of roots (the identity of the table, the targets of a search)
template <int NQ> __global__ __launch_bounds__(64) void k_root_keys(RootKeyArgs a) {
  constexpr int P = 1 << (2 * NQ);
  const int lane = threadIdx.x;
  for (long long i = blockIdx.x; i < a.n; i += gridDim.x) {
    uint64_t lo = 0, hi = 0;
    if (lane < P)
      for (int r = 0; r < P; r++) {
        const uint32_t k = a.R.rv[r], v = a.root_fp[(size_t)i * P * P + r * P + lane];
        lo += (uint64_t)(k & 0xffffu) * v;
        hi += (uint64_t)(k >> 16) * v;
      }
    const uint32_t h = m31_fold(lo, hi);
    const uint64_t s = wsum64(lane < P ? col_term(m31_red((uint64_t)h * h)) : 0ull);
    if (lane == 0) a.keys[i] = coset_key(s);
  }
}

struct KeysArgs {
  Rot R;
  Forest F;
  int t;                       // level of the parents
  const uint32_t *plist;       // parents (node indices at level t); null: every node
  long long np;
  uint64_t *keys;              // by child index k * (P-1) + p - 1
  uint32_t *k32, *v32;         // sort prefix and child index
};

// every child's coset key (src/tcount.cu:197-207): one wave per parent, the parent's F_p column in registers,
// each child column hash a linear form of it
template <int NQ> __global__ __launch_bounds__(64) void k_keys(KeysArgs a) {
  constexpr int P = 1 << (2 * NQ), G = P - 1;
  __shared__ uint32_t M[P * P];
  __shared__ uint8_t seq[MAXT];
  const int lane = threadIdx.x;
  for (long long k = blockIdx.x; k < a.np; k += gridDim.x) {
    const uint32_t node = a.plist ? a.plist[k] : (uint32_t)k;
    const uint32_t *src = M;
    if (a.F.mat[a.t]) src = a.F.mat[a.t] + (size_t)node * (P * P);
    else load_node<P>(a.F, a.R, a.t, node, M, seq, lane);
    const uint32_t inst = inst_of(a.F, a.t, node);
    uint32_t v[P];
#pragma unroll
    for (int r = 0; r < P; r++) v[r] = lane < P ? src[r * P + lane] : 0u;
    uint64_t mine = 0;
    for (int p = 1; p <= G; p++) {
      const uint32_t *C = a.R.coef + (p - 1) * P;
      uint64_t lo = 0, hi = 0;
#pragma unroll
      for (int r = 0; r < P; r++) {
        const uint32_t c = C[r];
        lo += (uint64_t)(c & 0xffffu) * v[r];
        hi += (uint64_t)(c >> 16) * v[r];
      }
      const uint32_t h = m31_fold(lo, hi);
      const uint64_t s = wsum64(lane < P ? col_term(m31_red((uint64_t)h * h)) : 0ull);
      if (lane == p - 1) mine = s;
    }
    if (lane < G) {
      const uint64_t key = coset_key(mine);
      const long long ch = k * G + lane;
      a.keys[ch] = key;
      a.k32[ch] = sort_prefix(key, inst);
      a.v32[ch] = (uint32_t)ch;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

struct DedupeArgs {
  Rot R;
  Forest F;
  int t;
  const uint32_t *plist;
  const uint32_t *k32, *v32;
  long long n;
  const uint32_t *runs;
  const unsigned long long *nruns;
  const uint64_t *keys;
  uint8_t *keep;               // by child; cleared for every child whose coset an earlier child holds
  int *flags;                  // bit 1: a run held more than DMAXREP distinct cosets
  unsigned long long *stat;    // [0] children visited in runs, [1] exact comparisons, [2] duplicates
};

// A child (pa, p) whose parent is pa = (gp, q) with p < q and P_p, P_q commuting is the same matrix as the child
// (pb, q) of pa's sibling pb = (gp, p): R(P_p) and R(P_q) commute.  pb precedes pa (siblings are ordered by
// rotation, between them at most q - p - 1 others), so that child has the smaller index and this one is a
// duplicate whenever pb is a node of the level (it may have been dropped as a duplicate or as a coset seen at a
// lower level).  One lane per candidate position of pb.
__device__ bool swap_duplicate(const Forest &F, const Rot &R, int t, uint32_t pa, int p, int lane) {
  if (t < 1) return false;
  const int q = F.rot[t][pa];
  if (p >= q || ((R.anti[p - 1] >> q) & 1)) return false;
  const uint32_t gp = F.par[t][pa];
  const long long i = (long long)pa - 1 - lane;
  const bool hit = lane < q - p && i >= 0 && F.par[t][i] == gp && F.rot[t][i] == p;
  return __ballot(hit) != 0;
}

// exact dedupe: within each run of equal prefixes (children in increasing index), the first child of every
// exact coset stays (src/tcount.cu:327-360: the smallest index of each coset).  Commuting swaps of an earlier
// child are marked without a comparison.
template <int NQ> __global__ __launch_bounds__(64) void k_dedupe(DedupeArgs a) {
  constexpr int P = 1 << (2 * NQ), G = P - 1;
  __shared__ uint32_t A[P * P], B[P * P];
  __shared__ uint8_t seq[MAXT];
  __shared__ uint32_t rep[DMAXREP];
  const int lane = threadIdx.x;
  bool full = false;
  unsigned long long n_vis = 0, n_cmp = 0, n_dup = 0;
  const long long nr = (long long)*a.nruns;
  auto parent = [&](uint32_t ch) { return a.plist ? a.plist[ch / G] : ch / G; };
  for (long long r = blockIdx.x; r < nr; r += gridDim.x) {
    const long long i0 = a.runs[r];
    const uint32_t kp = a.k32[i0];
    int nrep = 0;
    long long bid = -1;
    uint64_t kb = 0;
    for (long long j = i0; j < a.n && a.k32[j] == kp; j++) {
      const uint32_t ch = a.v32[j];
      const uint64_t key = a.keys[ch];
      const uint32_t inst = inst_of(a.F, a.t, parent(ch));
      bool dup = swap_duplicate(a.F, a.R, a.t, parent(ch), (int)(ch % G) + 1, lane), built = false;
      uint64_t ka = 0;
      for (int q = 0; q < nrep && !dup; q++) {
        const uint32_t rc = rep[q];
        if (a.keys[rc] != key || inst_of(a.F, a.t, parent(rc)) != inst) continue;
        if (!built) {
          ka = load_child<P>(a.F, a.R, a.t, parent(ch), (int)(ch % G) + 1, A, seq, lane);
          built = true;
        }
        if (bid != (long long)rc) {
          kb = load_child<P>(a.F, a.R, a.t, parent(rc), (int)(rc % G) + 1, B, seq, lane);
          bid = rc;
        }
        dup = same_cols<P>(A, ka, B, kb, lane);
        n_cmp++;
      }
      n_vis++;
      if (dup) {
        n_dup++;
        if (lane == 0) a.keep[ch] = 0;
      } else if (nrep < DMAXREP) {
        if (lane == 0) rep[nrep] = ch;
        nrep++;
        __builtin_amdgcn_wave_barrier();
      } else {
        full = true;
      }
    }
  }
  if (lane == 0 && full) atomicOr(a.flags, 2);
  if (lane == 0 && n_vis) {
    atomicAdd(&a.stat[0], n_vis);
    atomicAdd(&a.stat[1], n_cmp);
    atomicAdd(&a.stat[2], n_dup);
  }
}

struct MatArgs {
  Rot R;
  Forest F;                    // holds level t (par, rot) and a materialised level below it
  int t;
  long long n;
  uint32_t *out;               // n x P*P
};

// materialise the nodes of level t: each replayed from the nearest materialised level and written row-major
template <int NQ> __global__ __launch_bounds__(64) void k_mat(MatArgs a) {
  constexpr int P = 1 << (2 * NQ);
  __shared__ uint32_t M[P * P];
  __shared__ uint8_t seq[MAXT];
  const int lane = threadIdx.x;
  for (long long i = blockIdx.x; i < a.n; i += gridDim.x) {
    load_node<P>(a.F, a.R, a.t, (uint32_t)i, M, seq, lane);
    uint32_t *o = a.out + (size_t)i * (P * P);
#pragma unroll 16
    for (int k = lane; k < P * P; k += 64) o[k] = M[k];
    __builtin_amdgcn_wave_barrier();
  }
}

// HBM coset table: buckets of 8 slots (key | 1, level << 32 | node), probed by 8 lanes
struct HTab {
  unsigned long long *key, *val;
  unsigned long long bmask;
};

struct InsertArgs {
  HTab T;
  const uint64_t *keys;
  long long n;
  int level;
};

// insert the nodes of one level (no lookups run meanwhile); equal keys take separate slots
__global__ __launch_bounds__(256) void k_insert(InsertArgs a) {
  const int l8 = threadIdx.x & 7, shift = threadIdx.x & 56;
  for (long long i0 = (long long)blockIdx.x * 32; i0 < a.n; i0 += (long long)gridDim.x * 32) {
    const long long i = i0 + (threadIdx.x >> 3);
    bool done = i >= a.n;
    const uint64_t k = done ? 1ull : (a.keys[i] | 1ull);
    uint64_t b = tmix(k) & a.T.bmask;
    while (__ballot(!done)) {
      unsigned long long kv = 1;
      if (!done) kv = __hip_atomic_load(&a.T.key[b * 8 + l8], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t emp = (uint32_t)(__ballot(!done && kv == 0) >> shift) & 0xffu;
      unsigned long long old = 1;
      const int e = emp ? __ffs((int)emp) - 1 : 0;
      if (!done && emp && l8 == e) {
        old = atomicCAS(&a.T.key[b * 8 + e], 0ull, (unsigned long long)k);
        if (old == 0) a.T.val[b * 8 + e] = ((unsigned long long)a.level << 32) | (unsigned long long)i;
      }
      old = __shfl(old, shift + e, 64);
      if (!done) {
        if (emp && old == 0) done = true;
        else if (!emp) b = (b + 1) & a.T.bmask;
      }
    }
  }
}

struct ProbeArgs {
  HTab T;
  const uint64_t *keys;        // by candidate index
  const uint32_t *items;       // candidates to probe (null: 0 .. n-1)
  long long n;
  uint8_t *hit;                // by candidate index: 1 when some slot holds the same key
};

// key probe, 8 lanes per candidate
__global__ __launch_bounds__(256) void k_probe(ProbeArgs a) {
  const int l8 = threadIdx.x & 7, shift = threadIdx.x & 56;
  for (long long i0 = (long long)blockIdx.x * 32; i0 < a.n; i0 += (long long)gridDim.x * 32) {
    const long long i = i0 + (threadIdx.x >> 3);
    bool done = i >= a.n, hit = false;
    const uint32_t c = done ? 0u : (a.items ? a.items[i] : (uint32_t)i);
    const uint64_t k = done ? 1ull : (a.keys[c] | 1ull);
    uint64_t b = tmix(k) & a.T.bmask;
    while (__ballot(!done)) {
      const unsigned long long kv = done ? 1ull : a.T.key[b * 8 + l8];
      const uint32_t gm = (uint32_t)(__ballot(!done && kv == k) >> shift) & 0xffu;
      const uint32_t ge = (uint32_t)(__ballot(!done && kv == 0) >> shift) & 0xffu;
      if (!done) {
        if (gm) { hit = true; done = true; }
        else if (ge) done = true;
        else b = (b + 1) & a.T.bmask;
      }
    }
    if (i < a.n && l8 == 0) a.hit[c] = hit ? 1 : 0;
  }
}

struct ConfirmArgs {
  Rot R;
  HTab T;
  Forest Fc;                   // candidates' forest
  int tc;                      // mode 0: level of the candidates' parents; mode 1: level of the candidates
  const uint32_t *plist;       // mode 0: parents of the children
  Forest Ft;                   // the table's forest
  const uint32_t *items;       // candidates with a key hit
  long long n;
  const uint64_t *keys;        // by candidate index
  const int *maxc;             // mode 1: per instance, the largest table level accepted
  int mode;                    // 0: table filter, 1: search lookup
  uint8_t *fresh;              // mode 0: cleared when the child's coset is in the table
  unsigned long long *best;    // mode 1: per instance min(level << 32 | node)
  uint32_t *hitx;              // mode 1: table node of each hit, by node
  int *flags;
};

// exact confirmation of key hits (src/tcount.cu:248-276 lookup, :361-380 table filter): one wave per candidate
template <int NQ> __global__ __launch_bounds__(64) void k_confirm(ConfirmArgs a) {
  constexpr int P = 1 << (2 * NQ), G = P - 1;
  __shared__ uint32_t A[P * P], B[P * P];
  __shared__ uint8_t seq[MAXT];
  const int lane = threadIdx.x;
  for (long long it = blockIdx.x; it < a.n; it += gridDim.x) {
    const uint32_t c = a.items[it];
    const uint64_t k = a.keys[c] | 1ull;
    uint32_t node;
    int p;
    if (a.mode == 0) { node = a.plist ? a.plist[c / G] : c / G; p = (int)(c % G) + 1; }
    else { node = c; p = 0; }
    const uint32_t inst = inst_of(a.Fc, a.tc, node);
    const int maxc = a.maxc ? a.maxc[inst] : (1 << 30);
    int bestl = 1 << 30;
    uint32_t bestx = 0;
    bool built = false;
    uint64_t ka = 0;
    uint64_t b = tmix(k) & a.T.bmask;
    while (true) {
      const unsigned long long kv = lane < 8 ? a.T.key[b * 8 + lane] : 0ull;
      const unsigned long long vv = (lane < 8 && kv == k) ? a.T.val[b * 8 + lane] : 0ull;
      uint64_t mm = __ballot(lane < 8 && kv == k);
      const bool empty = __ballot(lane < 8 && kv == 0) != 0;
      while (mm) {
        const int s = __ffsll((long long)mm) - 1;
        mm &= mm - 1;
        const unsigned long long v = __shfl(vv, s, 64);
        const int lvl = (int)(v >> 32);
        const uint32_t x = (uint32_t)v;
        if (lvl > maxc || lvl >= bestl) continue;
        if (!built) {
          ka = load_child<P>(a.Fc, a.R, a.tc, node, p, A, seq, lane);
          built = true;
        }
        const uint64_t kb = load_child<P>(a.Ft, a.R, lvl, x, 0, B, seq, lane);
        if (same_cols<P>(A, ka, B, kb, lane)) { bestl = lvl; bestx = x; }
      }
      if (empty || (a.mode == 0 && bestl < (1 << 30))) break;
      b = (b + 1) & a.T.bmask;
    }
    if (lane == 0 && bestl < (1 << 30)) {
      if (a.mode == 0) a.fresh[c] = 0;
      else {
        a.hitx[c] = bestx;
        atomicMin(&a.best[inst], ((unsigned long long)bestl << 32) | (unsigned long long)c);
      }
    }
  }
}

// flag functors for the ordered compactions
struct FlagU8 {
  const uint8_t *f;
  __device__ bool operator()(long long i) const { return f[i] != 0; }
};
struct FlagActive {            // nodes whose instance is still searching
  const uint32_t *inst;
  const uint8_t *active;
  __device__ bool operator()(long long i) const { return active[inst ? inst[i] : 0] != 0; }
};

struct GatherArgs {
  const uint32_t *list;        // kept children, in increasing index
  long long n;
  const uint32_t *plist;       // parents (null: identity)
  const uint32_t *pinst;       // parents' instances (null: one instance)
  const uint64_t *ckeys;
  int G;
  uint32_t *par;
  uint8_t *rot;
  uint32_t *inst;              // may be null
  uint64_t *keys;
};

__global__ void k_gather(GatherArgs a) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (long long)gridDim.x * blockDim.x) {
    const uint32_t ch = a.list[i];
    const uint32_t par = a.plist ? a.plist[ch / a.G] : ch / a.G;
    a.par[i] = par;
    a.rot[i] = (uint8_t)(ch % a.G + 1);
    if (a.inst) a.inst[i] = a.pinst ? a.pinst[par] : 0u;
    a.keys[i] = a.ckeys[ch];
  }
}

__global__ void k_fill_u8(uint8_t *p, long long n, uint8_t v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, long long n, unsigned long long v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_copy_u8(uint8_t *d, const uint8_t *s, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) d[i] = s[i];
}

// first and one-past-last node of every instance in a level (nodes are ordered by instance)
__global__ void k_inst_bounds(const uint32_t *inst, long long n, long long *lo, long long *hi) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const uint32_t s = inst[i];
    if (i == 0 || inst[i - 1] != s) lo[s] = i;
    if (i == n - 1 || inst[i + 1] != s) hi[s] = i + 1;
  }
}

struct PathArgs {
  Forest Fs, Ft;
  int j;                       // search level of the hits
  long long n;
  const uint32_t *node;
  const uint32_t *hitx;        // table node matched by each search node of level j
  const int *tlevel;
  int *out;                    // n x 2 MAXT: search rotations (levels 1..j), then table rotations (levels 1..c)
};

__global__ void k_paths(PathArgs a) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < a.n; q += (long long)gridDim.x * blockDim.x) {
    int *o = a.out + q * 2 * MAXT;
    uint32_t i = a.node[q];
    for (int l = a.j; l >= 1; l--) { o[l - 1] = a.Fs.rot[l][i]; i = a.Fs.par[l][i]; }
    i = a.hitx[a.node[q]];
    for (int l = a.tlevel[q]; l >= 1; l--) { o[MAXT + l - 1] = a.Ft.rot[l][i]; i = a.Ft.par[l][i]; }
  }
}

// ------------------------------------------------------------------------------------------------------
// host side

using chz::DBuf;
using chz::set_err;

struct Spec {
  int n = 0, P = 0, e = 0;
  std::vector<int> x, y, tr, ts;
};

// .tspec reader (src/tcount.cu:433-450): 2 malformed, 3 n outside 1..3
// whitespace-separated tokens of a spec text, read in place (a batch parses hundreds of 64 x 64 channels)
struct Tok {
  const char *p, *end;
  void skip() { while (p < end && std::isspace((unsigned char)*p)) p++; }
  bool word(std::string &w) {
    skip();
    const char *b = p;
    while (p < end && !std::isspace((unsigned char)*p)) p++;
    w.assign(b, p);
    return p > b;
  }
  template <typename T> bool num(T &v) {
    skip();
    if (p >= end) return false;
    char *q = nullptr;
    const long long x = std::strtoll(p, &q, 10);
    if (q == p || q > end) return false;
    p = q;
    v = (T)x;
    return true;
  }
};

static int read_spec(const char *text, size_t len, Spec &s, std::string &err) {
  // strtoll stops at the terminating NUL of a std::string copy; texts from the C ABI need not end in one
  const std::string buf(text, len);
  Tok in{buf.c_str(), buf.c_str() + buf.size()};
  std::string tok;
  in.word(tok);
  if (tok != "tcount-spec") { err = "not a tcount spec"; return 2; }
  int ver = 0;
  in.num(ver);
  bool have_ch = false, have_rot = false;
  while (in.word(tok)) {
    if (tok == "n") {
      if (!in.num(s.n)) { err = "malformed n"; return 2; }
      if (s.n < 1 || s.n > 3) { err = "n = " + std::to_string(s.n) + " not supported (1 <= n <= 3)"; return 3; }
      s.P = 1 << (2 * s.n);
    } else if (tok == "channel") {
      if (!s.P) { err = "channel before n"; return 2; }
      bool ok = in.num(s.e);
      s.x.resize(s.P * s.P);
      s.y.resize(s.P * s.P);
      for (int i = 0; i < s.P * s.P && ok; i++) ok = in.num(s.x[i]) && in.num(s.y[i]);
      if (!ok) { err = "malformed channel"; return 2; }
      have_ch = true;
    } else if (tok == "rotations") {
      int m = 0;
      if (!in.num(m) || !s.P || m != s.P - 1) { err = "malformed rotations"; return 2; }
      s.tr.resize(m * s.P);
      s.ts.resize(m * s.P);
      bool ok = true;
      for (int i = 0; i < m * s.P && ok; i++) ok = in.num(s.tr[i]) && in.num(s.ts[i]);
      if (!ok) { err = "malformed rotations"; return 2; }
      have_rot = true;
    } else if (tok == "end") {
      break;
    } else {
      err = "unknown field " + tok;
      return 2;
    }
  }
  if (s.n < 1 || s.n > 3) { err = "n = " + std::to_string(s.n) + " not supported (1 <= n <= 3)"; return 3; }
  if (!have_ch || !have_rot) { err = "spec lacks channel or rotations"; return 2; }
  return 0;
}

// the Pauli multiplication tables the peel relies on: with digits 0 I, 1 X, 2 Y, 3 Z per qubit, P_p P_a is a
// multiple of P_(a ^ p), so an anticommuting row's partner is row a ^ p, which anticommutes too, and its sign is
// +-1 (tcount.py's writer produces exactly these)
static bool tables_ok(const Spec &s) {
  const int P = s.P;
  for (int p = 1; p < P; p++)
    for (int a = 0; a < P; a++) {
      const int r = s.tr[(p - 1) * P + a], sg = s.ts[(p - 1) * P + a];
      if (r != -1 && r != (a ^ p)) return false;
      if (r >= 0 && (s.tr[(p - 1) * P + r] != a || (sg != 1 && sg != -1))) return false;
    }
  return true;
}

enum { F_KEYS = 0, F_SORT, F_DEDUPE, F_PROBE, F_CONFIRM, F_COMPACT, F_INSERT, F_MAT, F_N };
static const char *f_names[F_N] = {"keys", "sort", "dedupe", "probe", "confirm", "compact", "insert", "materialise"};

// per-level device arrays of one forest; mat[l]: the level's nodes materialised (mat[0]: the roots, owned
// by the caller), kept only for the level being expanded and the one below it
struct Levels {
  std::vector<uint32_t *> par, inst, mat;
  std::vector<uint8_t *> rot;
  std::vector<uint64_t *> keys;
  std::vector<long long> count;
  void release() {
    (void)chz::sync_owner();           // the blocks return to the cache: no kernel may still use them
    for (auto p : par) chz::dev_free(p);
    for (auto p : inst) chz::dev_free(p);
    for (auto p : rot) chz::dev_free(p);
    for (auto p : keys) chz::dev_free(p);
    for (size_t l = 1; l < mat.size(); l++) chz::dev_free(mat[l]);
    par.clear(); inst.clear(); rot.clear(); keys.clear(); count.clear(); mat.clear();
  }
};

template <typename T> static T *dalloc(size_t n) {
  return (T *)chz::dev_alloc(std::max<size_t>(n, 1) * sizeof(T));
}

static unsigned grid1(long long n, long long cap = 65536) { return (unsigned)std::max<long long>(1, std::min(n, cap)); }
static unsigned gridt(long long n, int threads = 256, long long cap = 16384) {
  return (unsigned)std::max<long long>(1, std::min((n + threads - 1) / threads, cap));
}

}  // namespace tcn

struct chz_tc {
  int n = 0, P = 0, G = 0, K = 0, Kt = 0, device = 0;
  hipStream_t stream = nullptr;
  tcn::Spec spec;                            // the spec the table was built with (its rotation tables)
  uint64_t *d_anti = nullptr, *d_plus = nullptr;
  uint32_t *d_coef = nullptr, *d_rv = nullptr, *d_rw = nullptr;
  // table forest: one root, the identity
  tcn::Levels tab;
  uint32_t *d_troot_fp = nullptr;
  double mat_budget = 32e9;                  // bytes of materialised levels one forest may hold
  std::vector<chz_tc_level> tlog;
  // HBM table
  unsigned long long *d_hkey = nullptr, *d_hval = nullptr;
  unsigned long long nslots = 0;
  long long nentries = 0;
  // work arrays
  chz::DBuf<uint64_t> ckeys;
  chz::DBuf<uint32_t> sk[2], sv[2], runs, list, list2;
  chz::DBuf<uint8_t> keep, hit, fresh;
  chz::DBuf<unsigned long long> cnt, bcount, hist, tmp;
  chz::DBuf<int> flags;
  // measurement
  chz::EventLog ev;
  long k_launch[tcn::F_N] = {};
  double dcount[4] = {};                      // dedupe: runs, children visited, exact comparisons, duplicates
  double k_ms[tcn::F_N] = {}, k_work[tcn::F_N] = {};
};

namespace tcn {

static Rot rot_of(const chz_tc *c) { return Rot{c->d_anti, c->d_plus, c->d_coef, c->d_rv, c->d_rw}; }

static Forest forest_of(const Levels &L) {
  Forest F{};
  for (size_t l = 0; l < L.par.size() && l <= (size_t)MAXT; l++) {
    F.par[l] = L.par[l];
    F.rot[l] = L.rot[l];
    F.inst[l] = L.inst.size() > l ? L.inst[l] : nullptr;
    F.mat[l] = L.mat.size() > l ? L.mat[l] : nullptr;
  }
  return F;
}

// materialise level t of L (its parents must be reachable from a materialised level below) when it fits the
// budget, and drop the materialised level t-1, which nothing reads any more
template <int NQ> static void materialise(chz_tc *c, Levels &L, int t) {
  constexpr int P = 1 << (2 * NQ);
  while (L.mat.size() <= (size_t)t) L.mat.push_back(nullptr);
  if (t >= 1 && !L.mat[t] && (double)L.count[t] * P * P * 4 <= c->mat_budget && L.count[t] > 0) {
    uint32_t *m = dalloc<uint32_t>((size_t)L.count[t] * P * P);
    c->ev.begin(F_MAT, c->stream);
    MatArgs ma{rot_of(c), forest_of(L), t, L.count[t], m};
    hipLaunchKernelGGL(k_mat<NQ>, dim3(grid1(L.count[t])), dim3(64), 0, c->stream, ma);
    c->ev.end(c->stream);
    L.mat[t] = m;
  }
  if (t >= 2 && L.mat[t - 1] && L.mat[t]) {
    HIPCHK(hipStreamSynchronize(c->stream));
    chz::dev_free(L.mat[t - 1]);
    L.mat[t - 1] = nullptr;
  }
}

// ordered compaction of flagged indices 0..n-1 into out; returns the count
template <typename F> static long long compact(chz_tc *c, F f, long long n, uint32_t *out) {
  if (n == 0) return 0;
  const long long nb = (n + CT_TILE - 1) / CT_TILE;
  c->bcount.need(nb + 1);
  c->tmp.need(nb / SC_TILE + 4);
  hipLaunchKernelGGL(k_flag_count<F>, dim3((unsigned)nb), dim3(CT_THREADS), 0, c->stream, f, n, c->bcount.p);
  HIPCHK(scan_excl(c->bcount.p, nb, c->tmp.p, c->stream));
  unsigned long long tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, c->bcount.p + nb, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (tot && out)
    hipLaunchKernelGGL(k_scatter_idx<F>, dim3((unsigned)nb), dim3(CT_THREADS), 0, c->stream, f, n, c->bcount.p, out);
  return (long long)tot;
}

// sorted (prefix, child) pairs of nc children in sk[0]/sv[0]
static void sort_children(chz_tc *c, long long nc) {
  c->sk[1].need(nc);
  c->sv[1].need(nc);
  const long long nb = (nc + RS_TILE - 1) / RS_TILE;
  c->hist.need(nb * 256 + 1);
  c->tmp.need(nb * 256 / SC_TILE + 4);
  HIPCHK(radix_sort_pairs(c->sk[0].p, c->sv[0].p, c->sk[1].p, c->sv[1].p, nc, c->hist.p, c->tmp.p, c->stream));
}

struct GrowOut {
  long long kept = 0, fresh = 0;
  int flags = 0;
};

// one level of growth of forest L (the table when table = true) from the nodes plist of level t (null: all np).
// Children are keyed, sorted, deduplicated exactly, (table) filtered against the HBM table, and the survivors
// become level t+1 (src/tcount.cu:327-395).
template <int NQ>
static GrowOut grow(chz_tc *c, Levels &L, int t, const uint32_t *plist, long long np, bool table) {
  constexpr int P = 1 << (2 * NQ), G = P - 1;
  GrowOut out;
  const long long nc = np * G;
  if (nc >= (1ll << 32) - 1) { set_err("a level holds more than 2^32 children"); throw chz::DeviceError(); }
  materialise<NQ>(c, L, t);
  const Rot R = rot_of(c);
  const Forest F = forest_of(L);
  c->ckeys.need(nc);
  c->sk[0].need(nc);
  c->sv[0].need(nc);
  c->keep.need(nc);
  c->flags.need(1);
  HIPCHK(hipMemsetAsync(c->flags.p, 0, sizeof(int), c->stream));

  c->ev.begin(F_KEYS, c->stream);
  KeysArgs ka{R, F, t, plist, np, c->ckeys.p, c->sk[0].p, c->sv[0].p};
  hipLaunchKernelGGL(k_keys<NQ>, dim3(grid1(np)), dim3(64), 0, c->stream, ka);
  c->ev.end(c->stream);
  c->k_work[F_KEYS] += (double)nc * P * 2;

  c->ev.begin(F_SORT, c->stream);
  sort_children(c, nc);
  c->ev.end(c->stream);
  c->k_work[F_SORT] += (double)nc * 8 * 4 * 2;

  c->ev.begin(F_DEDUPE, c->stream);
  c->runs.need(nc);
  c->cnt.need(4);
  HIPCHK(hipMemsetAsync(c->cnt.p, 0, 32, c->stream));
  hipLaunchKernelGGL(k_fill_u8, dim3(gridt(nc)), dim3(256), 0, c->stream, c->keep.p, nc, (uint8_t)1);
  hipLaunchKernelGGL(k_run_starts, dim3(grid_for(nc / (64 * RH_ITEMS) + 1)), dim3(256), 0, c->stream, c->sk[0].p, nc,
                     c->runs.p, c->cnt.p, nullptr);
  unsigned long long nruns = 0;
  HIPCHK(hipMemcpyAsync(&nruns, c->cnt.p, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (nruns) {
    DedupeArgs da{R, F, t, plist, c->sk[0].p, c->sv[0].p, nc, c->runs.p, c->cnt.p, c->ckeys.p, c->keep.p, c->flags.p,
                  c->cnt.p + 1};
    hipLaunchKernelGGL(k_dedupe<NQ>, dim3(grid1((long long)nruns)), dim3(64), 0, c->stream, da);
  }
  c->ev.end(c->stream);

  c->list.need(nc);
  c->ev.begin(F_COMPACT, c->stream);
  out.kept = compact(c, FlagU8{c->keep.p}, nc, c->list.p);
  c->ev.end(c->stream);
  if (table && out.kept) {
    c->hit.need(nc);
    c->fresh.need(nc);
    c->ev.begin(F_PROBE, c->stream);
    hipLaunchKernelGGL(k_fill_u8, dim3(gridt(nc)), dim3(256), 0, c->stream, c->hit.p, nc, (uint8_t)0);
    ProbeArgs pa{HTab{c->d_hkey, c->d_hval, c->nslots / 8 - 1}, c->ckeys.p, c->list.p, out.kept, c->hit.p};
    hipLaunchKernelGGL(k_probe, dim3(gridt(out.kept, 32)), dim3(256), 0, c->stream, pa);
    hipLaunchKernelGGL(k_copy_u8, dim3(gridt(nc)), dim3(256), 0, c->stream, c->fresh.p, c->keep.p, nc);
    c->ev.end(c->stream);
    c->list2.need(nc);
    const long long nh = compact(c, FlagU8{c->hit.p}, nc, c->list2.p);
    if (nh) {
      c->ev.begin(F_CONFIRM, c->stream);
      ConfirmArgs ca{R, HTab{c->d_hkey, c->d_hval, c->nslots / 8 - 1}, F, t, plist, forest_of(c->tab), c->list2.p,
                     nh, c->ckeys.p, nullptr, 0, c->fresh.p, nullptr, nullptr, c->flags.p};
      hipLaunchKernelGGL(k_confirm<NQ>, dim3(grid1(nh)), dim3(64), 0, c->stream, ca);
      c->ev.end(c->stream);
    }
    c->ev.begin(F_COMPACT, c->stream);
    out.fresh = compact(c, FlagU8{c->fresh.p}, nc, c->list.p);
    c->ev.end(c->stream);
  } else {
    out.fresh = out.kept;
  }

  // level t+1
  const long long nn = out.fresh;
  uint32_t *par = dalloc<uint32_t>(nn);
  uint8_t *rot = dalloc<uint8_t>(nn);
  uint64_t *keys = dalloc<uint64_t>(nn);
  uint32_t *inst = table ? nullptr : dalloc<uint32_t>(nn);
  L.par.push_back(par);
  L.rot.push_back(rot);
  L.keys.push_back(keys);
  L.inst.push_back(inst);
  L.count.push_back(nn);
  L.mat.resize(L.par.size(), nullptr);
  if (nn) {
    c->ev.begin(F_COMPACT, c->stream);
    GatherArgs ga{c->list.p, nn, plist, L.inst[t], c->ckeys.p, G, par, rot, inst, keys};
    hipLaunchKernelGGL(k_gather, dim3(gridt(nn)), dim3(256), 0, c->stream, ga);
    c->ev.end(c->stream);
  }
  HIPCHK(hipMemcpyAsync(&out.flags, c->flags.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  unsigned long long st[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(st, c->cnt.p, 32, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->ev.flush(c->k_ms, c->k_launch);
  c->k_work[F_DEDUPE] += (double)st[2];
  c->dcount[0] += (double)st[0];
  c->dcount[1] += (double)st[1];
  c->dcount[2] += (double)st[2];
  c->dcount[3] += (double)st[3];
  return out;
}

// (re)build the HBM table for `need` entries from every level's keys
static void table_reserve(chz_tc *c, long long need) {
  if (c->d_hkey && (double)need <= 0.6 * (double)c->nslots) return;
  unsigned long long slots = 1ull << 12;
  while ((double)need > 0.45 * (double)slots) slots <<= 1;
  if (c->d_hkey) {
    (void)hipStreamSynchronize(c->stream);
    chz::dev_free(c->d_hkey);
    chz::dev_free(c->d_hval);
    c->d_hkey = c->d_hval = nullptr;
  }
  c->d_hkey = dalloc<unsigned long long>(slots);
  c->d_hval = dalloc<unsigned long long>(slots);
  c->nslots = slots;
  hipLaunchKernelGGL(k_fill_u64, dim3(gridt((long long)slots)), dim3(256), 0, c->stream, c->d_hkey, (long long)slots, 0ull);
  c->nentries = 0;
  for (size_t l = 0; l < c->tab.keys.size(); l++) {
    const long long n = c->tab.count[l];
    if (!n) continue;
    InsertArgs ia{HTab{c->d_hkey, c->d_hval, slots / 8 - 1}, c->tab.keys[l], n, (int)l};
    hipLaunchKernelGGL(k_insert, dim3(gridt(n, 32)), dim3(256), 0, c->stream, ia);
    c->nentries += n;
  }
}

static void table_insert_level(chz_tc *c, int l) {
  const long long n = c->tab.count[l];
  if (!c->d_hkey || (double)(c->nentries + n) > 0.6 * (double)c->nslots) {
    table_reserve(c, c->nentries + n + (c->nentries + n) / 2);
    return;   // the rebuild inserted every level, this one included
  }
  if (!n) return;
  c->ev.begin(F_INSERT, c->stream);
  InsertArgs ia{HTab{c->d_hkey, c->d_hval, c->nslots / 8 - 1}, c->tab.keys[l], n, l};
  hipLaunchKernelGGL(k_insert, dim3(gridt(n, 32)), dim3(256), 0, c->stream, ia);
  c->ev.end(c->stream);
  c->nentries += n;
}

static void upload_rotations(chz_tc *c) {
  const int P = c->P, G = c->G;
  std::vector<uint64_t> anti(G, 0), plus(G, 0);
  for (int p = 0; p < G; p++)
    for (int a = 0; a < P; a++) {
      if (c->spec.tr[p * P + a] >= 0) anti[p] |= 1ull << a;
      if (c->spec.ts[p * P + a] > 0) plus[p] |= 1ull << a;
    }
  // column hash weights R[r] (a fixed pseudo-random sequence) and per-rotation child coefficients: the child's
  // value is v'_a = v_a (commuting row a) or (v_a - s_a v_b) / sqrt2 (b its partner), so its column hash
  // sum_a R[a] v'_a equals sum_a coef[p][a] v_a with coef = R[a] or (R[a] - s_b R[b]) / sqrt2
  std::vector<uint32_t> rv(P), rw(P), coef(G * P);
  uint64_t s = 0x5eed7c0ull;
  for (int r = 0; r < P; r++) {
    s += 0x9e3779b97f4a7c15ull;
    rv[r] = (uint32_t)(tmix(s) % MP);
    s += 0x9e3779b97f4a7c15ull;
    rw[r] = (uint32_t)(tmix(s) % MP);
  }
  for (int p = 0; p < G; p++)
    for (int a = 0; a < P; a++) {
      const int b = c->spec.tr[p * P + a];
      if (b < 0) coef[p * P + a] = rv[a];
      else {
        const int sb = c->spec.ts[p * P + b];
        coef[p * P + a] = m31_rot(sb > 0 ? m31_sub(rv[a], rv[b]) : m31_add(rv[a], rv[b]), 15);
      }
    }
  c->d_anti = dalloc<uint64_t>(G);
  c->d_plus = dalloc<uint64_t>(G);
  c->d_coef = dalloc<uint32_t>(G * P);
  c->d_rv = dalloc<uint32_t>(P);
  c->d_rw = dalloc<uint32_t>(P);
  chz::copy_sync(c->d_anti, anti.data(), G * 8, hipMemcpyHostToDevice, c->stream);
  chz::copy_sync(c->d_plus, plus.data(), G * 8, hipMemcpyHostToDevice, c->stream);
  chz::copy_sync(c->d_coef, coef.data(), G * P * 4, hipMemcpyHostToDevice, c->stream);
  chz::copy_sync(c->d_rv, rv.data(), P * 4, hipMemcpyHostToDevice, c->stream);
  chz::copy_sync(c->d_rw, rw.data(), P * 4, hipMemcpyHostToDevice, c->stream);
}

// roots as residues
static void root_residues(const Spec &sp, uint32_t *pf) {
  for (size_t i = 0; i < sp.x.size(); i++) pf[i] = m31_entry(sp.x[i], sp.y[i], sp.e);
}

// the channel matrix is orthogonal (sum_a M[a][b] M[a][c] = delta_bc exactly), which the exactness of the
// residues relies on; entries must lie within the bound sqrt2^e
static bool orthogonal(const Spec &sp) {
  const int P = sp.P;
  if (sp.e < 0 || sp.e > EMAX) return false;
  const double lim = std::ldexp(1.0, (sp.e + 1) / 2) + 1;
  for (int i = 0; i < P * P; i++)
    if (std::abs((double)sp.x[i]) > lim || std::abs((double)sp.y[i]) > lim) return false;
  for (int b = 0; b < P; b++)
    for (int c = b; c < P; c++) {
      long long A = 0, B = 0;
      for (int a = 0; a < P; a++) {
        const long long x1 = sp.x[a * P + b], y1 = sp.y[a * P + b], x2 = sp.x[a * P + c], y2 = sp.y[a * P + c];
        A += x1 * x2 + 2 * y1 * y2;
        B += x1 * y2 + y1 * x2;
      }
      if (B != 0 || A != (b == c ? (1ll << sp.e) : 0)) return false;
    }
  return true;
}

template <int NQ> static void root_keys(chz_tc *c, const uint32_t *root_fp, long long n, uint64_t *keys) {
  RootKeyArgs ra{rot_of(c), root_fp, n, keys};
  hipLaunchKernelGGL(k_root_keys<NQ>, dim3(grid1(n)), dim3(64), 0, c->stream, ra);
}

template <int NQ> static void build_table(chz_tc *c) {
  const int P = c->P;
  Spec id;
  id.n = c->n;
  id.P = P;
  id.x.assign(P * P, 0);
  id.y.assign(P * P, 0);
  for (int i = 0; i < P; i++) id.x[i * P + i] = 1;
  std::vector<uint32_t> hf(P * P);
  root_residues(id, hf.data());
  c->d_troot_fp = dalloc<uint32_t>(P * P);
  chz::copy_sync(c->d_troot_fp, hf.data(), P * P * 4, hipMemcpyHostToDevice, c->stream);
  Levels &T = c->tab;
  T.par.push_back(nullptr);
  T.rot.push_back(nullptr);
  T.inst.push_back(nullptr);
  T.mat.push_back(c->d_troot_fp);
  T.keys.push_back(dalloc<uint64_t>(1));
  T.count.push_back(1);
  root_keys<NQ>(c, c->d_troot_fp, 1, T.keys[0]);
  table_reserve(c, 1 << 10);
  c->Kt = 0;
  for (int l = 1; l <= c->K; l++) {
    const double t0 = chz::now_s();
    const long long np = T.count[l - 1];
    const GrowOut g = grow<NQ>(c, T, l - 1, nullptr, np, true);
    if (g.flags & 2) { set_err("a dedupe run holds more distinct cosets than its buffer"); throw chz::DeviceError(); }
    table_insert_level(c, l);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ev.flush(c->k_ms, c->k_launch);
    chz_tc_level lv{l, (long)np, (long)(np * c->G), (long)g.kept, (long)g.fresh, chz::now_s() - t0};
    c->tlog.push_back(lv);
    c->Kt = l;
    if (!g.fresh) break;
  }
  // the table's nodes are replayed from the identity when a lookup needs them; drop materialised levels
  HIPCHK(hipStreamSynchronize(c->stream));
  for (size_t l = 1; l < T.mat.size(); l++)
    if (T.mat[l]) { chz::dev_free(T.mat[l]); T.mat[l] = nullptr; }
}

// one batch of searches against the table (src/tcount.cu:497-531 per instance)
struct Inst {
  int slot;                    // index into the caller's arrays
  Spec s;
};

template <int NQ>
static void search_batch(chz_tc *c, std::vector<Inst> &ins, int maxt, chz_tc_result *res, double t0) {
  constexpr int P = 1 << (2 * NQ), G = P - 1;
  const int ni = (int)ins.size();
  if (!ni) return;
  Levels S;
  std::vector<void *> extra;
  struct Guard {
    Levels &S;
    std::vector<void *> &extra;
    ~Guard() { S.release(); for (void *p : extra) chz::dev_free(p); }     // S.release() synchronised
  } guard{S, extra};
  auto keep_ptr = [&](auto *p) { extra.push_back((void *)p); return p; };
  // roots
  std::vector<uint32_t> hf((size_t)ni * P * P), hinst(ni);
  for (int i = 0; i < ni; i++) {
    root_residues(ins[i].s, hf.data() + (size_t)i * P * P);
    hinst[i] = (uint32_t)i;
  }
  uint32_t *rf = keep_ptr(dalloc<uint32_t>(hf.size()));
  int *d_maxc = keep_ptr(dalloc<int>(ni));
  uint8_t *d_active = keep_ptr(dalloc<uint8_t>(ni));
  unsigned long long *d_best = keep_ptr(dalloc<unsigned long long>(ni));
  long long *d_lo = keep_ptr(dalloc<long long>(ni));
  long long *d_hi = keep_ptr(dalloc<long long>(ni));
  uint32_t *d_hitx = nullptr;
  chz::copy_sync(rf, hf.data(), hf.size() * 4, hipMemcpyHostToDevice, c->stream);
  S.par.push_back(nullptr);
  S.rot.push_back(nullptr);
  S.mat.push_back(rf);
  S.inst.push_back(dalloc<uint32_t>(ni));
  S.keys.push_back(dalloc<uint64_t>(ni));
  S.count.push_back(ni);
  chz::copy_sync(S.inst[0], hinst.data(), ni * 4, hipMemcpyHostToDevice, c->stream);
  root_keys<NQ>(c, rf, ni, S.keys[0]);

  std::vector<int> state(ni, -1);             // -1 searching, else the result code
  std::vector<long long> cnt_prev(ni, 1), cnt(ni, 1);
  std::vector<unsigned long long> best(ni);
  std::vector<uint8_t> active(ni, 1);
  const Rot R = rot_of(c);
  const Forest Ft = forest_of(c->tab);
  const HTab HT{c->d_hkey, c->d_hval, c->nslots / 8 - 1};
  int running = ni;
  for (int j = 0; j <= maxt && running; j++) {
    const double tl = chz::now_s();
    if (j > 0) {
      // parents: level j-1 nodes of instances still searching
      const long long np_all = S.count[j - 1];
      const uint32_t *plist = nullptr;
      long long np = np_all;
      bool all = true;
      for (int i = 0; i < ni; i++) all = all && active[i];
      if (!all) {
        HIPCHK(hipMemcpyAsync(d_active, active.data(), ni, hipMemcpyHostToDevice, c->stream));
        c->list2.need(np_all);
        np = compact(c, FlagActive{S.inst[j - 1], d_active}, np_all, c->list2.p);
        uint32_t *pl = keep_ptr(dalloc<uint32_t>(np));
        if (np) HIPCHK(hipMemcpyAsync(pl, c->list2.p, np * 4, hipMemcpyDeviceToDevice, c->stream));
        plist = pl;
      }
      const GrowOut g = grow<NQ>(c, S, j - 1, plist, np, false);
      if (g.flags & 2) { set_err("a dedupe run holds more distinct cosets than its buffer"); throw chz::DeviceError(); }
      // per-instance counts of the new level
      const long long nn = S.count[j];
      hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, c->stream, (unsigned long long *)d_lo, (long long)ni, 0ull);
      hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, c->stream, (unsigned long long *)d_hi, (long long)ni, 0ull);
      if (nn) hipLaunchKernelGGL(k_inst_bounds, dim3(gridt(nn)), dim3(256), 0, c->stream, S.inst[j], nn, d_lo, d_hi);
      std::vector<long long> lo(ni), hi(ni);
      HIPCHK(hipMemcpyAsync(lo.data(), d_lo, ni * 8, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(hi.data(), d_hi, ni * 8, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      const double secs = chz::now_s() - tl;
      for (int i = 0; i < ni; i++) {
        if (state[i] >= 0) continue;
        cnt[i] = hi[i] - lo[i];
        chz_tc_result &r = res[ins[i].slot];
        if (r.nlevels <= MAXT) {
          r.levels[r.nlevels] = chz_tc_level{j, (long)cnt_prev[i], (long)(cnt_prev[i] * G), (long)cnt[i], (long)cnt[i], secs};
          r.nlevels++;
        }
        if (!cnt[i]) state[i] = 1;   // nothing left to peel: T-count above maxt
      }
    }
    // lookup of level j
    const long long nn = S.count[j];
    std::vector<int> maxc(ni, std::min(c->Kt, maxt - j));
    HIPCHK(hipMemcpyAsync(d_maxc, maxc.data(), ni * 4, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, c->stream, d_best, (long long)ni, ~0ull);
    if (nn) {
      c->hit.need(nn);
      c->ev.begin(F_PROBE, c->stream);
      ProbeArgs pa{HT, S.keys[j], nullptr, nn, c->hit.p};
      hipLaunchKernelGGL(k_probe, dim3(gridt(nn, 32)), dim3(256), 0, c->stream, pa);
      c->ev.end(c->stream);
      c->list2.need(nn);
      const long long nh = compact(c, FlagU8{c->hit.p}, nn, c->list2.p);
      if (nh) {
        d_hitx = keep_ptr(dalloc<uint32_t>(nn));
        c->flags.need(1);
        HIPCHK(hipMemsetAsync(c->flags.p, 0, sizeof(int), c->stream));
        c->ev.begin(F_CONFIRM, c->stream);
        ConfirmArgs ca{R, HT, forest_of(S), j, nullptr, Ft, c->list2.p, nh, S.keys[j], d_maxc, 1, nullptr, d_best,
                       d_hitx, c->flags.p};
        hipLaunchKernelGGL(k_confirm<NQ>, dim3(grid1(nh)), dim3(64), 0, c->stream, ca);
        c->ev.end(c->stream);
      }
    }
    HIPCHK(hipMemcpyAsync(best.data(), d_best, ni * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ev.flush(c->k_ms, c->k_launch);
    // results of this level
    std::vector<uint32_t> qn;
    std::vector<int> ql, qi;
    for (int i = 0; i < ni; i++)
      if (state[i] < 0 && best[i] != ~0ull) {
        qn.push_back((uint32_t)best[i]);
        ql.push_back((int)(best[i] >> 32));
        qi.push_back(i);
      }
    if (!qn.empty()) {
      const long long nq = (long long)qn.size();
      uint32_t *dn = keep_ptr(dalloc<uint32_t>(nq));
      int *dl = keep_ptr(dalloc<int>(nq));
      int *dout = keep_ptr(dalloc<int>(nq * 2 * MAXT));
      HIPCHK(hipMemcpyAsync(dn, qn.data(), nq * 4, hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(dl, ql.data(), nq * 4, hipMemcpyHostToDevice, c->stream));
      PathArgs pa{forest_of(S), Ft, j, nq, dn, d_hitx, dl, dout};
      hipLaunchKernelGGL(k_paths, dim3(gridt(nq)), dim3(256), 0, c->stream, pa);
      std::vector<int> paths(nq * 2 * MAXT);
      HIPCHK(hipMemcpyAsync(paths.data(), dout, paths.size() * 4, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      for (long long q = 0; q < nq; q++) {
        const int i = qi[q], lv = ql[q];
        chz_tc_result &r = res[ins[i].slot];
        r.tcount = j + lv;
        r.split[0] = j;
        r.split[1] = lv;
        int k = 0;
        for (int l = 0; l < j; l++) { r.rotations[2 * k] = paths[q * 2 * MAXT + l]; r.rotations[2 * k + 1] = 0; k++; }
        for (int l = lv - 1; l >= 0; l--) { r.rotations[2 * k] = paths[q * 2 * MAXT + MAXT + l]; r.rotations[2 * k + 1] = 1; k++; }
        state[i] = 0;
      }
    }
    for (int i = 0; i < ni; i++) {
      if (state[i] < 0 && j + c->Kt >= maxt) state[i] = 1;
      if (state[i] >= 0 && active[i]) {
        active[i] = 0;
        running--;
        res[ins[i].slot].rc = state[i];
        res[ins[i].slot].seconds = chz::now_s() - t0;
      }
      cnt_prev[i] = cnt[i];
    }
  }
  for (int i = 0; i < ni; i++)
    if (state[i] < 0) { res[ins[i].slot].rc = 1; res[ins[i].slot].seconds = chz::now_s() - t0; }
}

}  // namespace tcn

// ------------------------------------------------------------------------------------------------------
// C ABI

extern "C" {

const char *chz_tc_last_error(void) { return chz::get_err(); }

int chz_tc_table(const char *spec_text, size_t len, int K, int device, chz_tc **out) {
  using namespace tcn;
  if (!spec_text || !out) { set_err("null argument"); return 3; }
  *out = nullptr;
  std::string err;
  Spec s;
  if (const int rc = read_spec(spec_text, len, s, err)) { set_err(err); return rc; }
  if (!tables_ok(s)) { set_err("rotation tables do not pair anticommuting rows"); return 3; }
  if (K < 0 || K + 1 > EMAX || K > MAXT) { set_err("table depth outside 0.." + std::to_string(EMAX - 1)); return 3; }
  chz_tc *c = new chz_tc();
  c->n = s.n;
  c->P = s.P;
  c->G = s.P - 1;
  c->K = K;
  c->device = device;
  c->spec = s;
  try {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    chz::StreamScope scope(c->stream);
    upload_rotations(c);
    switch (c->n) {
      case 1: build_table<1>(c); break;
      case 2: build_table<2>(c); break;
      default: build_table<3>(c); break;
    }
  } catch (const chz::DeviceError &) {
    const std::string e = chz::get_err();
    chz_tc_free(c);
    set_err(e);
    return e.find("buffer") != std::string::npos ? 4 : 5;
  }
  *out = c;
  return 0;
}

int chz_tc_table_levels(const chz_tc *c, chz_tc_level *lv, int max) {
  if (!c) return 0;
  const int n = (int)c->tlog.size();
  for (int i = 0; i < n && i < max && lv; i++) lv[i] = c->tlog[i];
  return n;
}

long long chz_tc_table_size(const chz_tc *c) {
  if (!c) return 0;
  long long s = 0;
  for (long long n : c->tab.count) s += n;
  return s;
}

int chz_tc_table_depth(const chz_tc *c) { return c ? c->Kt : 0; }
int chz_tc_num_qubits(const chz_tc *c) { return c ? c->n : 0; }

int chz_tc_search(chz_tc *c, const char *const *specs, const size_t *lens, int ninst, int maxt, chz_tc_result *res) {
  using namespace tcn;
  if (!c || ninst < 0 || (ninst > 0 && (!specs || !lens || !res))) { set_err("null argument"); return 3; }
  if (maxt < 0 || maxt > MAXT) { set_err("maxt outside 0..32"); return 3; }
  chz::StreamScope scope(c->stream);
  const double t0 = chz::now_s();
  // parse and check every target (exact orthogonality is O(P^3) each), spread over host threads
  std::vector<Inst> all(ninst);
  std::vector<int> rcs(ninst, 0);
  std::vector<std::string> errs(ninst);
  auto prep = [&](int i) {
    std::memset(&res[i], 0, sizeof(chz_tc_result));
    res[i].rc = -1;
    Inst &in = all[i];
    in.slot = i;
    std::string &err = errs[i];
    int rc = read_spec(specs[i], lens[i], in.s, err);
    if (!rc && (in.s.n != c->n || in.s.tr != c->spec.tr || in.s.ts != c->spec.ts)) {
      err = "spec's qubit count or rotation tables differ from the table's";
      rc = 3;
    }
    if (!rc && !orthogonal(in.s)) {
      err = "channel matrix is not orthogonal";
      rc = 2;
    }
    if (!rc && in.s.e + std::max(0, maxt - c->Kt) + 1 > EMAX) {
      err = "channel exponent plus search depth beyond " + std::to_string(EMAX);
      rc = 3;
    }
    rcs[i] = rc;
  };
  const int nth = std::max(1, std::min<int>({ninst / 8, (int)std::thread::hardware_concurrency(), 16}));
  if (nth <= 1) {
    for (int i = 0; i < ninst; i++) prep(i);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nth; t++)
      th.emplace_back([&, t]() { for (int i = t; i < ninst; i += nth) prep(i); });
    for (auto &x : th) x.join();
  }
  std::vector<Inst> ins;
  for (int i = 0; i < ninst; i++) {
    if (rcs[i]) { res[i].rc = rcs[i]; set_err(errs[i]); continue; }
    ins.push_back(std::move(all[i]));
  }
  try {
    HIPCHK(hipSetDevice(c->device));
    switch (c->n) {
      case 1: search_batch<1>(c, ins, maxt, res, t0); break;
      case 2: search_batch<2>(c, ins, maxt, res, t0); break;
      default: search_batch<3>(c, ins, maxt, res, t0); break;
    }
  } catch (const chz::DeviceError &) {
    return 5;
  }
  return 0;
}

int chz_tc_kernel_stats(const chz_tc *c, chz_kstat *out, int n) {
  if (!c) return 0;
  for (int f = 0; f < tcn::F_N && f < n; f++)
    if (out) out[f] = chz_kstat{tcn::f_names[f], c->k_launch[f], c->k_ms[f], c->k_work[f]};
  return tcn::F_N;
}

int chz_tc_dedupe_counts(const chz_tc *c, double *out) {
  if (!c || !out) return 0;
  for (int i = 0; i < 4; i++) out[i] = c->dcount[i];
  return 4;
}

void chz_tc_free(chz_tc *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  chz::StreamScope scope(c->stream);
  (void)chz::sync_owner();           // cached device blocks must be idle before they are reused
  c->tab.release();
  void *ps[] = {c->d_anti, c->d_plus, c->d_coef, c->d_rv, c->d_rw, c->d_troot_fp, c->d_hkey, c->d_hval};
  for (void *p : ps) chz::dev_free(p);
  c->ckeys.release();
  for (int i = 0; i < 2; i++) { c->sk[i].release(); c->sv[i].release(); }
  c->runs.release(); c->list.release(); c->list2.release();
  c->keep.release(); c->hit.release(); c->fresh.release();
  c->cnt.release(); c->bcount.release(); c->hist.release(); c->tmp.release(); c->flags.release();
  c->ev.destroy();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
