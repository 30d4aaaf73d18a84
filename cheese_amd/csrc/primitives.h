// Device primitives shared by the search, T-count and Clifford paths: ordered stream compaction (counts ->
// one-block scan -> scatter) and a stable LSD radix sort of (u32 key, u32 value) pairs with LDS-staged
// 8-bit digit ranking.  Kernels live in an unnamed namespace so that every translation unit has its own.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "dev.h"

namespace {
using namespace chz;

// ordered stream compaction: tiles of 256 threads x 16 items, counts -> one-block scan -> scatter
constexpr int CT_THREADS = 256, CT_ITEMS = 16, CT_TILE = CT_THREADS * CT_ITEMS;

// Kernels taking `nd` read their element count from device memory when it is not null (launches sized for a
// capacity, as in a captured graph); n is then the capacity the grid was sized for.
__device__ __forceinline__ long long dev_n(long long n, const unsigned long long *nd) {
  return nd ? (long long)*nd : n;
}

template <typename F>
__global__ __launch_bounds__(CT_THREADS) void k_flag_count(F f, long long n, unsigned long long *bcount,
                                                           const unsigned long long *nd = nullptr) {
  n = dev_n(n, nd);
  const long long t0 = (long long)blockIdx.x * CT_TILE;
  int c = 0;
  for (int i = 0; i < CT_ITEMS; i++) {
    const long long x = t0 + (long long)i * CT_THREADS + threadIdx.x;
    if (x < n && f(x)) c++;
  }
  __shared__ int ws[CT_THREADS / 64];
  c = gsum<64>(c);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int w = 0; w < CT_THREADS / 64; w++) s += ws[w];
    bcount[blockIdx.x] = s;
  }
}

// exclusive scan of v[0..n) in place by one block; v[n] receives the total
__global__ __launch_bounds__(1024) void k_scan_one_block(unsigned long long *v, long long n) {
  __shared__ unsigned long long ws[16];
  __shared__ unsigned long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (long long base = 0; base < n; base += 1024) {
    const long long i = base + threadIdx.x;
    const unsigned long long x = i < n ? v[i] : 0;
    unsigned long long inc = x;   // inclusive wave scan
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned long long y = __shfl_up(inc, d, 64);
      if (lane >= d) inc += y;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long s = carry;
      for (int k = 0; k < 16; k++) { const unsigned long long t = ws[k]; ws[k] = s; s += t; }
      carry = s;
    }
    __syncthreads();
    if (i < n) v[i] = ws[w] + inc - x;
    __syncthreads();
  }
  if (threadIdx.x == 0) v[n] = carry;
}

// block-wide exclusive scan of one flag per thread, in thread order; returns the block total
__device__ __forceinline__ int block_flag_scan(bool flag, int &excl, int *ws) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t b = __ballot(flag);
  const int below = __popcll(b & ((1ull << lane) - 1));
  if (lane == 0) ws[w] = __popcll(b);
  __syncthreads();
  int off = 0, tot = 0;
  for (int k = 0; k < CT_THREADS / 64; k++) { if (k < w) off += ws[k]; tot += ws[k]; }
  __syncthreads();
  excl = off + below;
  return tot;
}

// flagged positions -> ordered index list (kept survivor ranks, children needing the T bound)
template <typename F>
__global__ __launch_bounds__(CT_THREADS) void k_scatter_idx(F f, long long n, const unsigned long long *boff,
                                                            uint32_t *out, const unsigned long long *nd = nullptr) {
  n = dev_n(n, nd);
  __shared__ int ws[CT_THREADS / 64];
  const long long t0 = (long long)blockIdx.x * CT_TILE;
  long long pos = (long long)boff[blockIdx.x];
  for (int i = 0; i < CT_ITEMS; i++) {
    const long long x = t0 + (long long)i * CT_THREADS + threadIdx.x;
    const bool fl = x < n && f(x);
    int ex;
    const int tot = block_flag_scan(fl, ex, ws);
    if (fl) out[pos + ex] = (uint32_t)x;
    pos += tot;
  }
}

// one tile (n <= CT_TILE) in one launch: flagged positions -> ordered index list, total written to *tot
template <typename F>
__global__ __launch_bounds__(CT_THREADS) void k_compact_one(F f, long long n, uint32_t *out, unsigned long long *tot) {
  __shared__ int ws[CT_THREADS / 64];
  long long pos = 0;
  for (int i = 0; i < CT_ITEMS; i++) {
    const long long x = (long long)i * CT_THREADS + threadIdx.x;
    const bool fl = x < n && f(x);
    int ex;
    const int t = block_flag_scan(fl, ex, ws);
    if (fl) out[pos + ex] = (uint32_t)x;
    pos += t;
  }
  if (threadIdx.x == 0) *tot = (unsigned long long)pos;
}

// run starts (a position whose key differs from the one before and equals the one after), appended in tiles:
// each thread tests RH_ITEMS positions, a block scan places the tile's starts and one atomic per tile reserves
// their slots at *count (the runs' order does not matter: each is deduplicated on its own)
constexpr int RH_ITEMS = 16;
__attribute__((unused)) __global__ __launch_bounds__(256) void k_run_starts(const uint32_t *k, long long n,
                                                                            uint32_t *runs, unsigned long long *count,
                                                                            const unsigned long long *nd) {
  n = dev_n(n, nd);
  __shared__ unsigned long long s_wave[4], s_base;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long long tile = 256ll * RH_ITEMS;
  for (long long t0 = (long long)blockIdx.x * tile; t0 < n; t0 += (long long)gridDim.x * tile) {
    uint32_t mask = 0;
    for (int it = 0; it < RH_ITEMS; it++) {
      const long long i = t0 + it * 256ll + tid;
      if (i < n) {
        const uint32_t ki = k[i];
        const bool head = i == 0 || k[i - 1] != ki;
        if (head && i + 1 < n && k[i + 1] == ki) mask |= 1u << it;
      }
    }
    const uint32_t c = (uint32_t)__popc(mask);
    uint32_t x = c;                                  // inclusive scan of the starts over the wave
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, d);
      if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[w] = x;
    __syncthreads();
    if (tid == 0) {
      unsigned long long tot = 0;
      for (int j = 0; j < 4; j++) { const unsigned long long v = s_wave[j]; s_wave[j] = tot; tot += v; }
      s_base = tot ? atomicAdd(count, tot) : 0ull;
    }
    __syncthreads();
    unsigned long long pos = s_base + s_wave[w] + (x - c);
    for (int it = 0; it < RH_ITEMS; it++)
      if ((mask >> it) & 1u) runs[pos++] = (uint32_t)(t0 + it * 256ll + tid);
    __syncthreads();                                 // s_wave and s_base are reused by the next tile
  }
}

// ------------------------------------------------------------------------------------------------------
// small sorts in one launch: n <= SS_MAX (key prefix, rank) pairs as 64-bit composites (prefix << 32 | rank,
// all distinct) by a bitonic network in LDS.  Ascending composites are the stable order of the prefixes, so the
// result is the radix sort's.
constexpr int SS_MAX = 4096;
__attribute__((unused)) __global__ __launch_bounds__(1024) void k_sort_small(const uint64_t *key, long long n,
                                                                           uint32_t *kout, uint32_t *vout) {
  __shared__ uint64_t s[SS_MAX];
  if (n > SS_MAX) return;                  // callers take the radix sort beyond one block's LDS
  int np = 1;
  while (np < n) np <<= 1;
  for (int i = threadIdx.x; i < np; i += blockDim.x)
    s[i] = i < n ? ((key[i] >> 32) << 32) | (uint64_t)i : ~0ull;
  __syncthreads();
  for (int k = 2; k <= np; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < np; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const uint64_t a = s[i], b = s[l];
          if ((a > b) == ((i & k) == 0)) { s[i] = b; s[l] = a; }
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    kout[i] = (uint32_t)(s[i] >> 32);
    vout[i] = (uint32_t)s[i];
  }
}

// ------------------------------------------------------------------------------------------------------
// LSD radix sort of (u32 key, u32 value) pairs, 8-bit digits, stable: tiles of 256 threads x 8 rounds
constexpr int RS_THREADS = 256, RS_ROUNDS = 8, RS_TILE = RS_THREADS * RS_ROUNDS;

__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const uint32_t *keys, long long n, int shift,
                                                           unsigned long long *hist, long long nblocks,
                                                           const unsigned long long *nd = nullptr) {
  n = dev_n(n, nd);
  __shared__ unsigned int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long long t0 = (long long)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; r++) {
    const long long x = t0 + (long long)r * RS_THREADS + threadIdx.x;
    if (x < n) atomicAdd(&h[(keys[x] >> shift) & 255], 1u);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// this block's tile of (key, value) pairs placed from base[d], the first output position of digit d for the tile
// (shared; advanced as the rounds go)
__device__ __forceinline__ void radix_scatter_tile(const uint32_t *kin, const uint32_t *vin, long long n, int shift,
                                                   unsigned long long *base, uint32_t *kout, uint32_t *vout) {
  __shared__ unsigned int wc[RS_THREADS / 64][256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int k = 0; k < RS_THREADS / 64; k++) wc[k][tid] = 0;
  __syncthreads();
  const long long t0 = (long long)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; r++) {
    const long long x = t0 + (long long)r * RS_THREADS + tid;
    const bool valid = x < n;
    const uint32_t key = valid ? kin[x] : 0u;
    const uint32_t val = valid ? vin[x] : 0u;
    const int dg = (key >> shift) & 255;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const uint64_t bb = __ballot((dg >> b) & 1);
      peers &= ((dg >> b) & 1) ? bb : ~bb;
    }
    const int rank = __popcll(peers & ((1ull << lane) - 1));
    if (valid && rank == 0) wc[w][dg] = __popcll(peers);
    __syncthreads();
    unsigned int before = 0;
    if (valid) for (int k = 0; k < w; k++) before += wc[k][dg];
    if (valid) {
      const unsigned long long p = base[dg] + before + rank;
      kout[p] = key;
      vout[p] = val;
    }
    __syncthreads();
    unsigned int tot = 0;
    for (int k = 0; k < RS_THREADS / 64; k++) { tot += wc[k][tid]; wc[k][tid] = 0; }
    base[tid] += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(const uint32_t *kin, const uint32_t *vin, long long n,
                                                              int shift, const unsigned long long *hist,
                                                              long long nblocks, uint32_t *kout, uint32_t *vout,
                                                              const unsigned long long *nd = nullptr) {
  n = dev_n(n, nd);
  __shared__ unsigned long long base[256];
  base[threadIdx.x] = hist[(long long)threadIdx.x * nblocks + blockIdx.x];
  radix_scatter_tile(kin, vin, n, shift, base, kout, vout);
}

// Sorts of up to RS_DIRECT_MAX tiles skip the scan of the histogram: k_radix_count writes each tile's digit counts
// (cnt[tile * 256 + d], 32-bit) and every scatter block forms its own 256 start positions from them, the counts of
// all smaller digits plus those of its digit in earlier tiles.  Two launches per pass instead of six (count, three
// scan kernels, the scan's total, scatter); each block reads the whole count table, at most 512 KB from L2.
constexpr long long RS_DIRECT_MAX = 512;

__global__ __launch_bounds__(RS_THREADS) void k_radix_count(const uint32_t *keys, long long n, int shift,
                                                            uint32_t *cnt, const unsigned long long *nd = nullptr) {
  n = dev_n(n, nd);
  __shared__ unsigned int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long long t0 = (long long)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; r++) {
    const long long x = t0 + (long long)r * RS_THREADS + threadIdx.x;
    if (x < n) atomicAdd(&h[(keys[x] >> shift) & 255], 1u);
  }
  __syncthreads();
  cnt[(long long)blockIdx.x * 256 + threadIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter_direct(const uint32_t *kin, const uint32_t *vin,
                                                                     long long n, int shift, const uint32_t *cnt,
                                                                     long long nblocks, uint32_t *kout,
                                                                     uint32_t *vout,
                                                                     const unsigned long long *nd = nullptr) {
  n = dev_n(n, nd);
  __shared__ unsigned long long base[256];
  __shared__ unsigned long long wsum[RS_THREADS / 64];
  const int d = threadIdx.x, lane = d & 63, w = d >> 6;
  // digit d: its count over all tiles and over the tiles before this one (coalesced rows of 256 counts)
  unsigned long long tot = 0, below = 0;
  const long long me = blockIdx.x;
  long long t = 0;
  for (; t + 4 <= nblocks; t += 4) {
    const uint32_t c0 = cnt[t * 256 + d], c1 = cnt[(t + 1) * 256 + d], c2 = cnt[(t + 2) * 256 + d],
                   c3 = cnt[(t + 3) * 256 + d];
    tot += (unsigned long long)c0 + c1 + c2 + c3;
    below += (t < me ? c0 : 0u) + (t + 1 < me ? c1 : 0u) + (t + 2 < me ? c2 : 0u) + (t + 3 < me ? c3 : 0u);
  }
  for (; t < nblocks; t++) {
    const uint32_t c = cnt[t * 256 + d];
    tot += c;
    if (t < me) below += c;
  }
  // exclusive scan of the digit totals over the 256 digits
  unsigned long long inc = tot;
#pragma unroll
  for (int k = 1; k < 64; k <<= 1) {
    const unsigned long long y = __shfl_up(inc, k, 64);
    if (lane >= k) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  unsigned long long off = 0;
  for (int k = 0; k < w; k++) off += wsum[k];
  base[d] = off + inc - tot + below;
  __syncthreads();
  radix_scatter_tile(kin, vin, n, shift, base, kout, vout);
}


// ------------------------------------------------------------------------------------------------------
// exclusive scan of long arrays: tiles of 4096 entries scanned in place, tile totals scanned by one block,
// tile offsets added back
constexpr int SC_THREADS = 1024, SC_ITEMS = 4, SC_TILE = SC_THREADS * SC_ITEMS;

__global__ __launch_bounds__(SC_THREADS) void k_scan_tiles(unsigned long long *v, long long n,
                                                          unsigned long long *tot) {
  __shared__ unsigned long long ws[SC_THREADS / 64];
  const long long base = (long long)blockIdx.x * SC_TILE + (long long)threadIdx.x * SC_ITEMS;
  unsigned long long x[SC_ITEMS], s = 0;
#pragma unroll
  for (int i = 0; i < SC_ITEMS; i++) { x[i] = base + i < n ? v[base + i] : 0; s += x[i]; }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long inc = s;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0;
    for (int k = 0; k < SC_THREADS / 64; k++) { const unsigned long long t = ws[k]; ws[k] = a; a += t; }
    tot[blockIdx.x] = a;
  }
  __syncthreads();
  unsigned long long e = ws[w] + inc - s;
#pragma unroll
  for (int i = 0; i < SC_ITEMS; i++) {
    if (base + i < n) v[base + i] = e;
    e += x[i];
  }
}

__global__ void k_scan_add(unsigned long long *v, long long n, const unsigned long long *tot) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    v[i] += tot[i / SC_TILE];
}

// v[0..n) <- exclusive prefix sums, v[n] <- total.  tmp holds at least n / SC_TILE + 2 entries.
static inline hipError_t scan_excl(unsigned long long *v, long long n, unsigned long long *tmp, hipStream_t s) {
  if (n <= 4 * SC_TILE) {
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, s, v, n);
    return hipGetLastError();
  }
  const long long nt = (n + SC_TILE - 1) / SC_TILE;
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(SC_THREADS), 0, s, v, n, tmp);
  hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, s, tmp, nt);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)std::min<long long>((n + 255) / 256, 16384)), dim3(256), 0, s, v, n,
                     tmp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(v + n, tmp + nt, sizeof(unsigned long long), hipMemcpyDeviceToDevice, s);
}

// stable sort of n (u32 key, u32 value) pairs, 4 passes of 8 bits: (k0, v0) -> (k0, v0) through (k1, v1).
// hist holds at least 256 * ceil(n / RS_TILE) + 1 entries, tmp what scan_excl needs for that many.
// With nd the count is read on the device and n is the capacity the launches are sized for.
static inline hipError_t radix_sort_pairs(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, long long n,
                                          unsigned long long *hist, unsigned long long *tmp, hipStream_t s,
                                          const unsigned long long *nd = nullptr) {
  if (n <= 1) return hipSuccess;
  const long long nb = (n + RS_TILE - 1) / RS_TILE;
  uint32_t *ks[2] = {k0, k1}, *vs[2] = {v0, v1};
  if (nb <= RS_DIRECT_MAX) {
    uint32_t *cnt = reinterpret_cast<uint32_t *>(hist);     // nb * 256 counts fit the histogram's space
    for (int pass = 0; pass < 4; pass++) {
      const int a = pass & 1, b = a ^ 1;
      hipLaunchKernelGGL(k_radix_count, dim3((unsigned)nb), dim3(RS_THREADS), 0, s, ks[a], n, pass * 8, cnt, nd);
      hipLaunchKernelGGL(k_radix_scatter_direct, dim3((unsigned)nb), dim3(RS_THREADS), 0, s, ks[a], vs[a], n,
                         pass * 8, cnt, nb, ks[b], vs[b], nd);
    }
    return hipGetLastError();
  }
  for (int pass = 0; pass < 4; pass++) {
    const int a = pass & 1, b = a ^ 1;
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nb), dim3(RS_THREADS), 0, s, ks[a], n, pass * 8, hist, nb, nd);
    hipError_t e = scan_excl(hist, nb * 256, tmp, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nb), dim3(RS_THREADS), 0, s, ks[a], vs[a], n, pass * 8, hist,
                       nb, ks[b], vs[b], nd);
  }
  return hipGetLastError();
}

}  // namespace
