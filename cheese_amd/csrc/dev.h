// Device-side building blocks shared by the gate-count kernels: exact Z[w] entries held as int4
// (a, b, c, d = a + b w + c w^2 + d w^3, w = exp(i pi/4)), lane-group votes and shuffles for wave64,
// and Montgomery arithmetic for F_p ranks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chz {

constexpr uint32_t P_MOD = 998244353u;        // 119 * 2^23 + 1; w maps to 3^((p-1)/8)
constexpr uint32_t P_NINV = 998244351u;       // -p^-1 mod 2^32

struct DevGate {                               // S <- S g^-1 data of one candidate
  int nq, L, mask, ik;
  int q[3];
  int unit;                                    // every nonzero inverse entry is +-w^p
  int tdel;                                    // sqrt2-denominator exponent of the gate's Pauli transfer
                                               // matrix: a child's T bound is within tdel of its parent's
                                               // (0: Clifford, equal); 255: not bounded
  int cutloc;                                  // bit i: the gate acts on one side of cut i only
  int cdel;                                    // 3 bits per cut: a child's cut need is at least its parent's
                                               // minus this (ceil log2 of the gate's Schmidt rank); 7: none
  int cls;                                     // bits 0-7: index of cdel among the spec's distinct values
                                               // (KSpec::cdv); bits 8-15: 1 + the gate's T-bound class (0:
                                               // none): candidates of one class give children of equal T bound
  int4 inv[64];                                // L x L inverse local matrix, row m, column lc
  int8_t ph[64];                               // unit gates: w-power p of inv[m][lc], -1 for zero
};

constexpr int CXL_MAX = 128;                   // largest cx table held in the expansion kernel's LDS
constexpr int CD_MAX = 8;                      // distinct cut-delta classes cached per parent
constexpr int CSRC_MAX = 4;                    // cuts whose matrix-entry sources the expansion kernel stages
constexpr int TC_MAX = 8;                      // T-bound classes cached per parent

struct KSpec {                                 // device view of a spec (kernel argument)
  int n, G, depth, standard, lite, tdiv, gk, nsym, ncuts, have_table;
  int cx_mono;                                 // cx table non-decreasing in every cut need
  int cxlds;                                   // the cx table fits the expansion kernel's LDS copy (CXL_MAX bytes)
  int ncx;                                     // cx table entries
  int ncd;                                     // distinct gate cdel values (0: more than CD_MAX)
  uint32_t cdv[8];                             // those values; gate g's is cdv[gates[g].cls & 0xff]
  int cutmask[16], cutsize[16], rad[16];
  uint32_t wm[4];                              // Montgomery constants w^i * 2^32 mod p, i = 0..3
  const DevGate *gates;
  const uint8_t *bad2;                         // G x G
  const uint32_t *bad3;                        // bits over G^3
  const int4 *goal;                            // N x N
  uint64_t goalnz;                             // nonzero entries of the goal (bit r * N + c)
  const int *cxtable;
  const uint8_t *symperm;                      // nsym x N basis maps
  const uint8_t *symconj;                      // nsym
  uint32_t cjm;                                // bit y: symmetry y conjugates (y < 16)
  const uint4 *ktab;                           // the expansion kernel's LDS tables (LevelTabs), built on the host
  // goal pre-images: xg[g] (N x N, sqrt2 exponent xgk[g], nonzero-entry mask xgm[g]) is the one normalised
  // residual that candidate g maps to the goal (xgk -1: none); a parent whose coefficients are all at most
  // maglim has no child outside int32
  const int4 *xg;
  const uint64_t *xgm;
  const int *xgk;
  uint32_t maglim;
};

// Index of the calling wave in the grid (blocks of a multiple of 64 threads), read through readfirstlane: the
// compiler then treats it, and the loop bounds and indices derived from it, as wave-uniform scalars instead of
// per-lane vector values.
__device__ __forceinline__ long long wave_index() {
  return (long long)blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// ---- lane groups: GS lanes (a power of two dividing 64) cooperate on one matrix ----
template <int GS> __device__ __forceinline__ uint64_t gballot(bool p) {
  const uint64_t b = __ballot(p);
  if constexpr (GS == 64) return b;
  else return (b >> ((__lane_id()) & ~(GS - 1))) & ((1ull << GS) - 1);
}
template <int GS> __device__ __forceinline__ bool gall(bool p) {
  if constexpr (GS == 64) return gballot<GS>(p) == ~0ull;
  else return gballot<GS>(p) == ((1ull << GS) - 1);
}
template <int GS> __device__ __forceinline__ bool gany(bool p) { return gballot<GS>(p) != 0; }

// ---- whole-wave exchanges without the LDS pipeline (gfx950): DPP row controls and permlane16/32 swaps.  They
// need no per-lane address registers (ds_bpermute does), which keeps the expansion kernel's register pressure
// down.  Every lane of the wave must be active.
// DPP controls: quad permutations (lane ^ 1, ^ 2 within a quad), half-row mirror (lane ^ 7 within 8 lanes), row
// mirror (lane ^ 15 within a 16-lane row)
constexpr int DPP_QXOR1 = 0xB1, DPP_QXOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;
template <int CTRL> __device__ __forceinline__ uint32_t dpp32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}
// all-reduce over the wave of a commutative operation: each step pairs every lane with one of another group whose
// lanes already agree (quads, then 8, 16 and 32 lanes), so the pairing need not be a lane xor
template <typename Op> __device__ __forceinline__ uint32_t wave_reduce32(uint32_t v, Op op) {
  v = op(v, dpp32<DPP_QXOR1>(v));
  v = op(v, dpp32<DPP_QXOR2>(v));
  v = op(v, dpp32<DPP_HALF_MIRROR>(v));
  v = op(v, dpp32<DPP_MIRROR>(v));
  // v_permlane16_swap(v, v) returns rows (0, 0, 2, 2) and (1, 1, 3, 3) of v; v_permlane32_swap(v, v) the halves
  // (0, 0) and (1, 1)
  const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
  v = op(p[0], p[1]);
  const auto q = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return op(q[0], q[1]);
}

template <int GS> __device__ __forceinline__ int4 gshfl(int4 v, int src) {
  return make_int4(__shfl(v.x, src, GS), __shfl(v.y, src, GS), __shfl(v.z, src, GS), __shfl(v.w, src, GS));
}
template <int GS> __device__ __forceinline__ int4 gxor(int4 v, int m) {
  return make_int4(__shfl_xor(v.x, m, GS), __shfl_xor(v.y, m, GS), __shfl_xor(v.z, m, GS), __shfl_xor(v.w, m, GS));
}
template <int GS, typename T> __device__ __forceinline__ T gsum(T v) {
  if constexpr (GS == 64 && sizeof(T) == 4) {
    return (T)wave_reduce32((uint32_t)v, [](uint32_t a, uint32_t b) { return a + b; });
  } else {
#pragma unroll
    for (int m = 1; m < GS; m <<= 1) v += __shfl_xor(v, m, GS);
    return v;
  }
}
template <int GS, typename T> __device__ __forceinline__ T gor(T v) {
  if constexpr (GS == 64 && sizeof(T) == 4) {
    return (T)wave_reduce32((uint32_t)v, [](uint32_t a, uint32_t b) { return a | b; });
  } else if constexpr (GS == 64 && sizeof(T) == 8) {
    const auto o = [](uint32_t a, uint32_t b) { return a | b; };
    return (T)(((uint64_t)wave_reduce32((uint32_t)((uint64_t)v >> 32), o) << 32) | wave_reduce32((uint32_t)v, o));
  } else {
#pragma unroll
    for (int m = 1; m < GS; m <<= 1) v |= __shfl_xor(v, m, GS);
    return v;
  }
}

// ---- Z[w] entries ----
__device__ __forceinline__ bool zzero(int4 v) { return !(v.x | v.y | v.z | v.w); }
__device__ __forceinline__ bool zeq(int4 u, int4 v) { return u.x == v.x && u.y == v.y && u.z == v.z && u.w == v.w; }
__device__ __forceinline__ bool zhalf(int4 v) { return !((v.x ^ v.z) & 1) && !((v.y ^ v.w) & 1); }   // sqrt2 | v
// v / sqrt2 = ((b - d) / 2, (a + c) / 2, (b + d) / 2, (c - a) / 2), exact when zhalf (a = c, b = d mod 2).  In 32
// bits without overflow: with floor halves, (a + c) / 2 = (a >> 1) + (c >> 1) + (a & 1) and (c - a) / 2 =
// (c >> 1) - (a >> 1) when a and c have equal parity (the same as the int64 sums shifted, for every int32 input).
__device__ __forceinline__ int4 zdown(int4 v) {
  const int a = v.x >> 1, b = v.y >> 1, c = v.z >> 1, d = v.w >> 1;
  return make_int4(b - d, a + c + (v.x & 1), b + d + (v.y & 1), c - a);
}
__device__ __forceinline__ int4 zconj(int4 v) { return make_int4(v.x, -v.w, -v.z, -v.y); }

// ---- F_p, Montgomery form with R = 2^32 ----
__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t b) {   // a < 2^32, b < p -> a b / R mod p
  const uint64_t t = (uint64_t)a * b;
  const uint32_t m = (uint32_t)t * P_NINV;
  const uint64_t u = (t + (uint64_t)m * P_MOD) >> 32;
  const uint32_t r = (uint32_t)u;
  return r >= P_MOD ? r - P_MOD : r;
}
__device__ __forceinline__ uint32_t fp_sub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + P_MOD - b; }
__device__ __forceinline__ uint32_t fp_lift(int v) { return v < 0 ? (uint32_t)(v + 3 * (int64_t)P_MOD) : (uint32_t)v; }
// image of a + b w + c w^2 + d w^3 in F_p
__device__ __forceinline__ uint32_t fp_of(int4 v, const uint32_t *wm) {
  uint32_t s = mont_mul(fp_lift(v.x), wm[0]) + mont_mul(fp_lift(v.y), wm[1]);
  s = s >= P_MOD ? s - P_MOD : s;
  s += mont_mul(fp_lift(v.z), wm[2]);
  s = s >= P_MOD ? s - P_MOD : s;
  s += mont_mul(fp_lift(v.w), wm[3]);
  return s >= P_MOD ? s - P_MOD : s;
}

__device__ __forceinline__ int combine_bound(int h, int t, int c2, int supp) {   // src/cheese.cu:599-607
  int best = 1 << 30;
  const int hi = c2 > supp ? c2 : supp;
  for (int c = c2; c <= hi; c++) {
    const int one = (h + t > supp - 2 * c) ? h + t : supp - 2 * c;
    if (c + one < best) best = c + one;
  }
  return best;
}

// The largest t with combine_bound(h, t, c2, supp) <= rem (negative: none), for rem >= 0.  The bound is at most
// rem exactly when some c in [c2, max(c2, supp)] has c + h + t <= rem and supp - c <= rem; the smallest such c is
// max(c2, supp - rem), which never exceeds max(c2, supp).
__device__ __forceinline__ int t_budget(int h, int c2, int supp, int rem) {
  const int c = c2 > supp - rem ? c2 : supp - rem;
  return rem - h - c;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 32; x *= 0xd6e8feb86659fd93ull;
  x ^= x >> 32; x *= 0xd6e8feb86659fd93ull;
  return x ^ (x >> 32);
}

}  // namespace chz
