// The n-qubit Clifford group (n <= 3) tabulated on one MI355X: breadth-first search by gate count from the
// identity into an open-addressing HBM hash table (tableau -> minimal gate count), batched distance queries
// and optimal-word read-back.  Reference behaviour: src/clifford.cu:41-166 (tableaux, table, BFS),
// :207-210 (batched find), :251-283 (BFS driver, group-order check), :344-372 (read-back).
//
// Table layout: 2^lg 64-bit slots, slot = tableau << 8 | distance, 0 = empty (no tableau is 0).  Slots are
// grouped in buckets of 8 (one 64-byte sector); a group of 8 lanes probes a whole bucket per step, votes
// for a match or an empty slot with a ballot, and claims the first empty slot with a 64-bit CAS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cheese_hip.h"
#include "host.h"
#include "primitives.h"

// Device allocations of the Clifford table go through the process-wide capture lock: another thread (the request
// server's gate-count engines) may be capturing a level step, and hipFree synchronises the device.
static hipError_t cl_malloc(void **p, size_t bytes) {
  std::lock_guard<std::recursive_mutex> g(chz::capture_mu());
  return hipMalloc(p, bytes);
}
static hipError_t cl_free(void *p) {
  std::lock_guard<std::recursive_mutex> g(chz::capture_mu());
  return hipFree(p);
}

namespace chz_cl {

struct Gates { int n, ng; int k[32], a[32], b[32]; };   // gate set: H, S, Sdg, X, Z per qubit, then CX pairs

static Gates gate_set(int n) {
  Gates G{};
  G.n = n;
  for (int q = 0; q < n; q++)
    for (int k = 0; k < 5; k++) { G.k[G.ng] = k; G.a[G.ng] = q; G.b[G.ng] = 0; G.ng++; }
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++)
      if (a != b) { G.k[G.ng] = 5; G.a[G.ng] = a; G.b[G.ng] = b; G.ng++; }
  return G;
}

// Tableau: 2n rows (images of X_0..X_{n-1}, Z_0..Z_{n-1}), row = x | z << n | sign << 2n, for the Hermitian
// Pauli (-1)^sign i^(x.z) X^x Z^z.  Gate kinds: 0 H, 1 S, 2 Sdg, 3 X, 4 Z, 5 CX a->b, 6 Y.
__host__ __device__ inline uint32_t conj_row(uint32_t w, int n, int k, int a, int b) {
  const uint32_t nm = (1u << n) - 1;
  uint32_t x = w & nm, z = (w >> n) & nm, r = (w >> (2 * n)) & 1;
  const uint32_t xa = (x >> a) & 1, za = (z >> a) & 1;
  if (k == 0) { r ^= xa & za; x = (x & ~(1u << a)) | (za << a); z = (z & ~(1u << a)) | (xa << a); }
  else if (k == 1) { r ^= xa & za; z ^= xa << a; }
  else if (k == 2) { r ^= xa & (za ^ 1); z ^= xa << a; }
  else if (k == 3) { r ^= za; }
  else if (k == 4) { r ^= xa; }
  else if (k == 5) {
    const uint32_t xb = (x >> b) & 1, zb = (z >> b) & 1;
    r ^= xa & zb & (xb ^ za ^ 1);
    x ^= xa << b;
    z ^= zb << a;
  } else { r ^= xa ^ za; }
  return x | (z << n) | (r << (2 * n));
}
__host__ __device__ inline uint64_t apply(uint64_t T, int n, int k, int a, int b) {   // gate after the Clifford T
  const int rb = 2 * n + 1;
  const uint64_t rm = (1ull << rb) - 1;
  uint64_t out = 0;
  for (int i = 0; i < 2 * n; i++) out |= (uint64_t)conj_row((uint32_t)((T >> (i * rb)) & rm), n, k, a, b) << (i * rb);
  return out;
}
__host__ __device__ inline uint64_t identity(int n) {
  const int rb = 2 * n + 1;
  uint64_t T = 0;
  for (int i = 0; i < n; i++) { T |= (uint64_t)(1u << i) << (i * rb); T |= (uint64_t)(1u << (n + i)) << ((n + i) * rb); }
  return T;
}
__host__ __device__ inline int inv_kind(int k) { return k == 1 ? 2 : k == 2 ? 1 : k; }

__host__ __device__ inline uint64_t hmix(uint64_t x) {
  x ^= x >> 31; x *= 0x7fb5d329728ea185ull;
  x ^= x >> 27; x *= 0x81dadef4bc2dd44dull;
  return x ^ (x >> 33);
}

struct Table {
  unsigned long long *slot;
  uint64_t bmask;                  // buckets - 1
};

// group of 8 lanes (l8 = lane & 7) probes for key; insert: claim an empty slot with distance d.
// Returns the stored distance (>= 0), or -1 when absent (find) / -2 when this call inserted it.
template <bool INSERT> __device__ inline int probe8(Table T, uint64_t key, int d, int l8) {
  const int shift = __lane_id() & ~7;
  uint64_t b = hmix(key) & T.bmask;
  const unsigned long long want = (unsigned long long)((key << 8) | (uint64_t)d);
  while (true) {
    // while the table is being filled other CUs write it: read through to the coherent level so that a
    // failed claim is never followed by the same stale view of the bucket
    unsigned long long v = INSERT ? __hip_atomic_load(&T.slot[b * 8 + l8], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                  : T.slot[b * 8 + l8];
    const uint64_t hit = (__ballot((v >> 8) == key && v != 0) >> shift) & 0xff;
    if (hit) return (int)(__shfl(v, shift + __ffsll((long long)hit) - 1, 64) & 255);
    uint64_t emp = (__ballot(v == 0) >> shift) & 0xff;
    if (!emp) { b = (b + 1) & T.bmask; continue; }
    if (!INSERT) return -1;
    // the lowest empty slot: its lane tries to claim it, the group learns the outcome
    const int e = __ffsll((long long)emp) - 1;
    unsigned long long old = 0;
    if (l8 == e) old = atomicCAS(&T.slot[b * 8 + e], 0ull, want);
    old = __shfl(old, shift + e, 64);
    if (old == 0) return -2;
    if ((old >> 8) == key) return (int)(old & 255);
    // another key took the slot: probe the same bucket again
  }
}

struct BfsArgs {
  Gates G;
  Table T;
  const uint64_t *front;
  long long nf;
  uint64_t *next;
  unsigned long long *cnt;          // [0] new elements, [1] capacity overflow
  long long cap;
  int d;
};

constexpr int STAGE = 2048;

// one BFS level: every (frontier element, gate) pair, 8 lanes per candidate; new tableaux are staged in LDS
// and written out with one reservation per block
__global__ __launch_bounds__(256) void k_bfs(BfsArgs a) {
  __shared__ uint64_t buf[STAGE];
  __shared__ unsigned int nbuf;
  __shared__ unsigned long long base;
  if (threadIdx.x == 0) nbuf = 0;
  __syncthreads();
  const int l8 = threadIdx.x & 7;
  const long long total = a.nf * a.G.ng;
  const long long groups = (long long)gridDim.x * (blockDim.x / 8);
  const long long g0 = (long long)blockIdx.x * (blockDim.x / 8) + (threadIdx.x >> 3);
  for (long long r0 = 0; r0 < total; r0 += groups) {
    const long long c = r0 + g0;
    if (c < total) {
      const long long i = c / a.G.ng;
      const int g = (int)(c % a.G.ng);
      const uint64_t t = apply(a.front[i], a.G.n, a.G.k[g], a.G.a[g], a.G.b[g]);
      if (probe8<true>(a.T, t, a.d, l8) == -2 && l8 == 0) buf[atomicAdd(&nbuf, 1u)] = t;
    }
    __syncthreads();
    if (nbuf > STAGE - 256 / 8 || r0 + groups >= total) {
      if (threadIdx.x == 0) base = atomicAdd(&a.cnt[0], (unsigned long long)nbuf);
      __syncthreads();
      for (unsigned int j = threadIdx.x; j < nbuf; j += blockDim.x) {
        const unsigned long long p = base + j;
        if ((long long)p < a.cap) a.next[p] = buf[j];
        else a.cnt[1] = 1;
      }
      __syncthreads();
      if (threadIdx.x == 0) nbuf = 0;
      __syncthreads();
    }
  }
}

// tableaux of words (letters are gate-set indices, applied in order), one lane per word
__global__ void k_words(Gates G, const uint8_t *letters, const long long *offs, long long nq, uint64_t *tab) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += (long long)gridDim.x * blockDim.x) {
    uint64_t T = identity(G.n);
    for (long long j = offs[q]; j < offs[q + 1]; j++) {
      const int g = letters[j];
      T = apply(T, G.n, G.k[g], G.a[g], G.b[g]);
    }
    tab[q] = T;
  }
}

// batched lookups: 8 lanes per query (src/clifford.cu:207-210)
__global__ __launch_bounds__(256) void k_find(Table T, const uint64_t *q, long long nq, int *out) {
  const int l8 = threadIdx.x & 7;
  const long long groups = (long long)gridDim.x * (blockDim.x / 8);
  for (long long i0 = 0; i0 < nq; i0 += groups) {       // uniform trip count: every lane takes part in the ballots
    const long long i = i0 + (long long)blockIdx.x * (blockDim.x / 8) + (threadIdx.x >> 3);
    if (i < nq) {
      const int d = probe8<false>(T, q[i], 0, l8);
      if (l8 == 0) out[i] = d;
    }
  }
}

// optimal words: one wave per query; each step, lane g tests the predecessor through gate g (its distance
// must be one less) and the first such gate in gate order is taken (src/clifford.cu:354-372)
__global__ __launch_bounds__(64) void k_readback(Gates G, Table T, const uint64_t *q, long long nq, int maxlen,
                                                 uint8_t *words, int *lens) {
  const int lane = threadIdx.x;
  for (long long i = blockIdx.x; i < nq; i += gridDim.x) {
    uint64_t X = q[i];
    // distance of the query: lanes 0..7 probe cooperatively
    int d = -1;
    {
      const int dd = lane < 8 ? probe8<false>(T, X, 0, lane) : 0;
      d = __shfl(dd, 0, 64);
    }
    if (d < 0 || d > maxlen) { if (lane == 0) lens[i] = -1; continue; }
    if (lane == 0) lens[i] = d;
    for (int s = d; s > 0; s--) {
      uint64_t pre = 0;
      bool ok = false;
      if (lane < G.ng) {
        pre = apply(X, G.n, inv_kind(G.k[lane]), G.a[lane], G.b[lane]);
        // scalar linear probing of one key per lane over the same bucket layout
        uint64_t b = hmix(pre) & T.bmask;
        for (int guard = 0; guard < (1 << 20); guard++) {
          bool done = false;
          for (int j = 0; j < 8; j++) {
            const unsigned long long v = T.slot[b * 8 + j];
            if (!v) { done = true; break; }
            if ((v >> 8) == pre) { ok = (int)(v & 255) == s - 1; done = true; break; }
          }
          if (done) break;
          b = (b + 1) & T.bmask;
        }
      }
      const uint64_t bal = __ballot(ok);
      const int g = bal ? __ffsll((long long)bal) - 1 : -1;
      if (g < 0) { if (lane == 0) lens[i] = -1; break; }
      X = __shfl(pre, g, 64);
      if (lane == 0) words[i * maxlen + (s - 1)] = (uint8_t)g;
    }
  }
}

// elements in a canonical order (sorted by tableau), so that seeded samples do not depend on where concurrent
// inserts happened to place keys: occupied slots are compacted, then stably sorted by the low and then the high
// 32 bits of the tableau
struct FlagSlot {
  const unsigned long long *s;
  __device__ bool operator()(long long i) const { return s[i] != 0; }
};
__global__ void k_key_lo(const unsigned long long *slot, const uint32_t *idx, long long n, uint32_t *k, uint32_t *v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    k[i] = (uint32_t)(slot[idx[i]] >> 8);
    v[i] = (uint32_t)i;
  }
}
__global__ void k_key_hi(const unsigned long long *slot, const uint32_t *idx, const uint32_t *perm, long long n,
                         uint32_t *k, uint32_t *v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    k[i] = (uint32_t)(slot[idx[perm[i]]] >> 40);
    v[i] = perm[i];
  }
}
__global__ void k_key_final(const unsigned long long *slot, const uint32_t *idx, const uint32_t *perm, long long n,
                            uint64_t *out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = slot[idx[perm[i]]] >> 8;
}

// seeded uniform elements: splitmix64 indices into the sorted element list
__global__ void k_sample(const uint64_t *sorted, long long order, unsigned long long seed, long long n, uint64_t *out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    out[i] = sorted[z % (uint64_t)order];
  }
}

// order-independent digest: sum over entries of hmix-free mix of (tableau * 64 + distance), and the count
__global__ __launch_bounds__(256) void k_digest(Table T, long long nslots, unsigned long long *out) {
  unsigned long long s = 0, c = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += (long long)gridDim.x * blockDim.x) {
    const unsigned long long v = T.slot[i];
    if (v) { uint64_t x = (v >> 8) * 64 + (v & 255); x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33;
             x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33; s += x; c++; }
  }
  for (int m = 1; m < 64; m <<= 1) { s += __shfl_xor(s, m, 64); c += __shfl_xor(c, m, 64); }
  if ((threadIdx.x & 63) == 0) { atomicAdd(&out[0], s); atomicAdd(&out[1], c); }
}

static thread_local std::string g_err;

#include "cliffopt_kernels.h"

}  // namespace chz_cl

namespace chz_cl {

// Fixed staging for chz_cliff_query_words_noalloc: the tableaux of up to kQueryScratch words, written by k_words
// and read by k_find.  Each GPU the library runs on holds its own copy until the process ends.  Host code never
// uses this name as an address: scratch_address() resolves each device's copy once, with that device current.
constexpr long long kQueryScratch = 1ll << 20;
__device__ __align__(256) uint64_t scratch[kQueryScratch];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint64_t *g_scratch_dev[kMaxDevices];

// the device address of this device's scratch (the device must be current); 0 or 5 with g_err set
static int scratch_address(int device, uint64_t **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) { g_err = "device ordinal out of range"; return 3; }
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) { g_err = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e); return 5; }
    g_scratch_dev[device] = (uint64_t *)p;
  }
  *out = g_scratch_dev[device];
  return 0;
}

}  // namespace chz_cl

struct chz_cliff {
  int n = 0, device = 0;
  uint64_t *scratch = nullptr;         // this device's chz_cl::scratch (its device address)
  chz_cl::Gates G{};
  chz_cl::Table T{};
  long long nslots = 0;
  std::vector<long> levels;
  hipStream_t stream = nullptr;
  double build_ms = 0;
  double probes = 0;
  uint64_t *sorted = nullptr;          // elements sorted by tableau (built on the first sample)
  long long order = 0;
};

#define CLCHK(x)                                                                         \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) { chz_cl::g_err = std::string("HIP error: ") + hipGetErrorString(e_); return 5; } \
  } while (0)

extern "C" {

const char *chz_cliff_last_error(void) { return chz_cl::g_err.c_str(); }

int chz_cliff_build(int n, int device, chz_cliff **out) {
  chz::RelaxedCapture relaxed;
  *out = nullptr;
  if (n < 1 || n > 3) { chz_cl::g_err = "n must be 1..3"; return 3; }
  chz_cliff *c = new chz_cliff();
  c->n = n;
  c->device = device;
  c->G = chz_cl::gate_set(n);
  const long long order = n == 1 ? 24 : n == 2 ? 11520 : 92897280;   // 2^(n^2+2n) prod_j (4^j - 1)
  int lg = 6;
  while ((1ll << lg) < order * 4 / 3) lg++;
  c->nslots = 1ll << lg;
  c->T.bmask = (uint64_t)(c->nslots / 8 - 1);
  CLCHK(hipSetDevice(device));
  if (const int rc = chz_cl::scratch_address(device, &c->scratch)) { delete c; return rc; }
  CLCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  CLCHK(cl_malloc((void **)&c->T.slot, c->nslots * 8));
  CLCHK(hipMemsetAsync(c->T.slot, 0, c->nslots * 8, c->stream));
  const long long cap = std::max(1024ll, order / 2);
  uint64_t *fa = nullptr, *fb = nullptr;
  unsigned long long *cnt = nullptr;
  CLCHK(cl_malloc((void **)&fa, cap * 8));
  CLCHK(cl_malloc((void **)&fb, cap * 8));
  CLCHK(cl_malloc((void **)&cnt, 16));
  hipEvent_t e0, e1;
  CLCHK(hipEventCreate(&e0));
  CLCHK(hipEventCreate(&e1));
  CLCHK(hipEventRecord(e0, c->stream));
  // the identity takes its slot with distance 0
  {
    const uint64_t id = chz_cl::identity(n);
    const unsigned long long w = (unsigned long long)(id << 8);
    uint64_t b = chz_cl::hmix(id) & c->T.bmask;
    CLCHK(hipMemcpyAsync(c->T.slot + b * 8, &w, 8, hipMemcpyHostToDevice, c->stream));
    CLCHK(hipMemcpyAsync(fa, &id, 8, hipMemcpyHostToDevice, c->stream));
  }
  long long nf = 1, total = 1;
  c->levels.assign(1, 1);
  for (int d = 1; nf; d++) {
    CLCHK(hipMemsetAsync(cnt, 0, 16, c->stream));
    chz_cl::BfsArgs a{c->G, c->T, fa, nf, fb, cnt, cap, d};
    const long long groups = nf * c->G.ng;
    const int blocks = (int)std::min<long long>(std::max<long long>((groups + 31) / 32, 1), 8192);
    hipLaunchKernelGGL(chz_cl::k_bfs, dim3(blocks), dim3(256), 0, c->stream, a);
    unsigned long long h[2];
    CLCHK(hipMemcpyAsync(h, cnt, 16, hipMemcpyDeviceToHost, c->stream));
    CLCHK(hipStreamSynchronize(c->stream));
    if (h[1]) { chz_cl::g_err = "frontier buffer too small"; return 4; }
    c->probes += (double)groups;
    nf = (long long)h[0];
    total += nf;
    if (nf) c->levels.push_back((long)nf);
    std::swap(fa, fb);
  }
  CLCHK(hipEventRecord(e1, c->stream));
  CLCHK(hipEventSynchronize(e1));
  float ms = 0;
  CLCHK(hipEventElapsedTime(&ms, e0, e1));
  c->build_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)cl_free(fa);
  (void)cl_free(fb);
  (void)cl_free(cnt);
  *out = c;
  if (total != order) { chz_cl::g_err = "group order mismatch"; return 4; }   // src/clifford.cu:279-280
  return 0;
}

int chz_cliff_levels(const chz_cliff *c, long *counts, int maxd) {
  const int nl = (int)c->levels.size();
  for (int i = 0; i < nl && i < maxd; i++) counts[i] = c->levels[i];
  return nl;
}

int chz_cliff_num_gates(const chz_cliff *c) { return c->G.ng; }

double chz_cliff_build_ms(const chz_cliff *c) { return c->build_ms; }

int chz_cliff_digest(const chz_cliff *c, uint64_t *digest, long long *entries) {
  unsigned long long *d = nullptr, h[2] = {0, 0};
  CLCHK(hipSetDevice(c->device));
  CLCHK(cl_malloc((void **)&d, 16));
  CLCHK(hipMemsetAsync(d, 0, 16, c->stream));
  hipLaunchKernelGGL(chz_cl::k_digest, dim3(2048), dim3(256), 0, c->stream, c->T, c->nslots, d);
  CLCHK(hipMemcpyAsync(h, d, 16, hipMemcpyDeviceToHost, c->stream));
  CLCHK(hipStreamSynchronize(c->stream));
  (void)cl_free(d);
  *digest = h[0];
  *entries = (long long)h[1];
  return 0;
}

// the sorted element list, built once
static int cliff_sorted(chz_cliff *c) {
  if (c->sorted) return 0;
  const long long ns = c->nslots;
  const long long nb = (ns + CT_TILE - 1) / CT_TILE;
  unsigned long long *bc = nullptr, *tmp = nullptr, *hist = nullptr;
  uint32_t *idx = nullptr, *k0 = nullptr, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr;
  CLCHK(cl_malloc((void **)&bc, (nb + 1) * 8));
  CLCHK(cl_malloc((void **)&tmp, (ns / SC_TILE + 8) * 8));
  chz_cl::FlagSlot f{c->T.slot};
  hipLaunchKernelGGL(k_flag_count<chz_cl::FlagSlot>, dim3((unsigned)nb), dim3(CT_THREADS), 0, c->stream, f, ns, bc);
  CLCHK(scan_excl(bc, nb, tmp, c->stream));
  unsigned long long n = 0;
  CLCHK(hipMemcpyAsync(&n, bc + nb, 8, hipMemcpyDeviceToHost, c->stream));
  CLCHK(hipStreamSynchronize(c->stream));
  const long long rb = ((long long)n + RS_TILE - 1) / RS_TILE;
  CLCHK(cl_malloc((void **)&idx, std::max<unsigned long long>(n, 1) * 4));
  CLCHK(cl_malloc((void **)&k0, std::max<unsigned long long>(n, 1) * 4));
  CLCHK(cl_malloc((void **)&v0, std::max<unsigned long long>(n, 1) * 4));
  CLCHK(cl_malloc((void **)&k1, std::max<unsigned long long>(n, 1) * 4));
  CLCHK(cl_malloc((void **)&v1, std::max<unsigned long long>(n, 1) * 4));
  CLCHK(cl_malloc((void **)&hist, (rb * 256 + 1) * 8));
  CLCHK(cl_malloc((void **)&c->sorted, std::max<unsigned long long>(n, 1) * 8));
  hipLaunchKernelGGL(k_scatter_idx<chz_cl::FlagSlot>, dim3((unsigned)nb), dim3(CT_THREADS), 0, c->stream, f, ns, bc, idx);
  const unsigned g = (unsigned)std::min<long long>(((long long)n + 255) / 256 + 1, 16384);
  hipLaunchKernelGGL(chz_cl::k_key_lo, dim3(g), dim3(256), 0, c->stream, c->T.slot, idx, (long long)n, k0, v0);
  CLCHK(radix_sort_pairs(k0, v0, k1, v1, (long long)n, hist, tmp, c->stream));
  hipLaunchKernelGGL(chz_cl::k_key_hi, dim3(g), dim3(256), 0, c->stream, c->T.slot, idx, v0, (long long)n, k1, v1);
  CLCHK(radix_sort_pairs(k1, v1, k0, v0, (long long)n, hist, tmp, c->stream));
  hipLaunchKernelGGL(chz_cl::k_key_final, dim3(g), dim3(256), 0, c->stream, c->T.slot, idx, v1, (long long)n, c->sorted);
  CLCHK(hipStreamSynchronize(c->stream));
  c->order = (long long)n;
  void *fr[] = {bc, tmp, hist, idx, k0, v0, k1, v1};
  for (void *p : fr) (void)cl_free(p);
  return 0;
}

// count uniformly random group elements (tableaux), seeded: uniform indices into the elements sorted by tableau
int chz_cliff_sample(chz_cliff *c, uint64_t seed, long long count, uint64_t *tab) {
  CLCHK(hipSetDevice(c->device));
  if (const int rc = cliff_sorted(c)) return rc;
  if (c->order <= 0) { chz_cl::g_err = "empty table"; return 4; }
  uint64_t *d = nullptr;
  CLCHK(cl_malloc((void **)&d, std::max<long long>(count, 1) * 8));
  hipLaunchKernelGGL(chz_cl::k_sample, dim3((unsigned)std::min<long long>((count + 255) / 256 + 1, 4096)), dim3(256), 0,
                     c->stream, c->sorted, c->order, (unsigned long long)seed, count, d);
  CLCHK(hipMemcpyAsync(tab, d, count * 8, hipMemcpyDeviceToHost, c->stream));
  CLCHK(hipStreamSynchronize(c->stream));
  (void)cl_free(d);
  return 0;
}

// distances of nq words: letters[offs[i] .. offs[i+1]) are gate-set indices applied in order.
// ms (may be NULL) receives the device time of the word and lookup kernels.
int chz_cliff_query_words(chz_cliff *c, const uint8_t *letters, const long long *offs, long long nq, int *dist,
                          double *ms) {
  chz::RelaxedCapture relaxed;
  CLCHK(hipSetDevice(c->device));
  const long long nl = offs[nq];
  uint8_t *dl = nullptr;
  long long *doff = nullptr;
  uint64_t *tab = nullptr;
  int *dd = nullptr;
  CLCHK(cl_malloc((void **)&dl, std::max<long long>(nl, 1)));
  CLCHK(cl_malloc((void **)&doff, (nq + 1) * 8));
  CLCHK(cl_malloc((void **)&tab, std::max<long long>(nq, 1) * 8));
  CLCHK(cl_malloc((void **)&dd, std::max<long long>(nq, 1) * 4));
  CLCHK(hipMemcpyAsync(dl, letters, nl, hipMemcpyHostToDevice, c->stream));
  CLCHK(hipMemcpyAsync(doff, offs, (nq + 1) * 8, hipMemcpyHostToDevice, c->stream));
  hipEvent_t e0, e1;
  CLCHK(hipEventCreate(&e0));
  CLCHK(hipEventCreate(&e1));
  CLCHK(hipEventRecord(e0, c->stream));
  hipLaunchKernelGGL(chz_cl::k_words, dim3((unsigned)std::min<long long>((nq + 255) / 256 + 1, 4096)), dim3(256), 0,
                     c->stream, c->G, dl, doff, nq, tab);
  hipLaunchKernelGGL(chz_cl::k_find, dim3((unsigned)std::min<long long>((nq + 31) / 32 + 1, 8192)), dim3(256), 0,
                     c->stream, c->T, tab, nq, dd);
  CLCHK(hipEventRecord(e1, c->stream));
  CLCHK(hipMemcpyAsync(dist, dd, nq * 4, hipMemcpyDeviceToHost, c->stream));
  CLCHK(hipStreamSynchronize(c->stream));
  float t = 0;
  CLCHK(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)cl_free(dl);
  (void)cl_free(doff);
  (void)cl_free(tab);
  (void)cl_free(dd);
  return 0;
}

// chz_cliff_query_words with device buffers and no allocation: the tableaux are staged in this device's fixed
// scratch (at most kQueryScratch words per call); the kernels go on `stream` (NULL: the table's own stream,
// synchronised before returning).  Calls that overlap on one GPU share the scratch and race.
int chz_cliff_query_words_noalloc(chz_cliff *c, const uint8_t *d_letters, const long long *d_offs, long long nq,
                                  int *d_dist, void *stream) {
  if (nq < 0 || (nq > 0 && (!d_letters || !d_offs || !d_dist))) { chz_cl::g_err = "bad arguments"; return 3; }
  if (nq > chz_cl::kQueryScratch) {
    chz_cl::g_err = "chz_cliff_query_words_noalloc: " + std::to_string(nq) + " words exceed the scratch capacity of " +
                    std::to_string(chz_cl::kQueryScratch);
    return 4;
  }
  if (nq == 0) return 0;
  chz::RelaxedCapture relaxed;
  CLCHK(hipSetDevice(c->device));
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  hipLaunchKernelGGL(chz_cl::k_words, dim3((unsigned)std::min<long long>((nq + 255) / 256 + 1, 4096)), dim3(256), 0,
                     st, c->G, d_letters, d_offs, nq, c->scratch);
  hipLaunchKernelGGL(chz_cl::k_find, dim3((unsigned)std::min<long long>((nq + 31) / 32 + 1, 8192)), dim3(256), 0,
                     st, c->T, c->scratch, nq, d_dist);
  CLCHK(hipGetLastError());
  if (!stream) CLCHK(hipStreamSynchronize(st));
  return 0;
}

// tableau of one word (host arithmetic, same algebra as the device)
uint64_t chz_cliff_word_tableau(const chz_cliff *c, const uint8_t *letters, int len) {
  uint64_t T = chz_cl::identity(c->n);
  for (int i = 0; i < len; i++) {
    const int g = letters[i];
    T = chz_cl::apply(T, c->n, c->G.k[g], c->G.a[g], c->G.b[g]);
  }
  return T;
}

// optimal words of nq tableaux (gate-set indices, <= maxlen each); lens[i] = length or -1
int chz_cliff_readback(chz_cliff *c, const uint64_t *tab, long long nq, int maxlen, uint8_t *words, int *lens) {
  chz::RelaxedCapture relaxed;
  CLCHK(hipSetDevice(c->device));
  uint64_t *dt = nullptr;
  uint8_t *dw = nullptr;
  int *dlen = nullptr;
  CLCHK(cl_malloc((void **)&dt, std::max<long long>(nq, 1) * 8));
  CLCHK(cl_malloc((void **)&dw, std::max<long long>(nq * maxlen, 1)));
  CLCHK(cl_malloc((void **)&dlen, std::max<long long>(nq, 1) * 4));
  CLCHK(hipMemcpyAsync(dt, tab, nq * 8, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(chz_cl::k_readback, dim3((unsigned)std::min<long long>(std::max<long long>(nq, 1), 8192)),
                     dim3(64), 0, c->stream, c->G, c->T, dt, nq, maxlen, dw, dlen);
  CLCHK(hipMemcpyAsync(words, dw, nq * maxlen, hipMemcpyDeviceToHost, c->stream));
  CLCHK(hipMemcpyAsync(lens, dlen, nq * 4, hipMemcpyDeviceToHost, c->stream));
  CLCHK(hipStreamSynchronize(c->stream));
  (void)cl_free(dt);
  (void)cl_free(dw);
  (void)cl_free(dlen);
  return 0;
}

#include "cliffopt.h"

void chz_cliff_free(chz_cliff *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->T.slot) (void)cl_free(c->T.slot);
  if (c->sorted) (void)cl_free(c->sorted);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
