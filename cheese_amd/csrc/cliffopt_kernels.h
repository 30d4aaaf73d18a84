// Device side of the T-placement beam search (`cliffopt`, src/clifford.cu:168-206): tableau composition and
// inversion, the placements of a T gate, one beam step with an in-block top-B selection, and the final segment
// lookup.  Included by clifford.hip inside namespace chz_cl, after the tableau and table definitions.
//
// A T gate on qubit q commutes with every Clifford K with K Z_q' K^-1 = +-Z_q, so T_q can be replaced by
// K T'_q' K^-1 (T' with its dagger flipped for the minus sign) at the cost of a global phase; segment j of the
// circuit then becomes K_{j+1}^-1 D_j K_j.  The beam search picks K for every T gate so that the summed optimal
// segment lengths (read from the Clifford table) are small; T = Tdg S and Tdg = T Sdg exactly, so a T may also
// flip its dagger, moving an S (Sdg) into the segment before it.

__host__ __device__ inline uint32_t trow(uint64_t T, int n, int i) {
  const int rb = 2 * n + 1;
  return (uint32_t)((T >> (i * rb)) & ((1ull << rb) - 1));
}
__host__ __device__ inline uint64_t tset(uint64_t T, int n, int i, uint32_t r) {
  const int rb = 2 * n + 1;
  const uint64_t m = ((1ull << rb) - 1) << (i * rb);
  return (T & ~m) | ((uint64_t)r << (i * rb));
}

// the Clifford's image of the Hermitian Pauli i^(x.z) X^x Z^z, as a row: the product of the images of its X and
// then Z factors, the power of i tracked through (X^a Z^b)(X^c Z^d) = (-1)^(b.c) X^(a^c) Z^(b^d)
__host__ __device__ inline uint32_t timage(uint64_t T, int n, uint32_t x, uint32_t z) {
  const uint32_t nm = (1u << n) - 1;
  int k = __builtin_popcount(x & z);
  uint32_t ax = 0, az = 0;
  for (int g = 0; g < 2 * n; g++) {
    const bool on = g < n ? ((x >> g) & 1) : ((z >> (g - n)) & 1);
    if (!on) continue;
    const uint32_t w = trow(T, n, g), bx = w & nm, bz = (w >> n) & nm, r = (w >> (2 * n)) & 1;
    k += 2 * (int)r + __builtin_popcount(bx & bz) + 2 * __builtin_popcount(az & bx);
    ax ^= bx;
    az ^= bz;
  }
  const uint32_t sign = (uint32_t)(((k - __builtin_popcount(ax & az)) & 3) >> 1);
  return ax | (az << n) | (sign << (2 * n));
}

// A then B: row i of the product is B's image of A's row i, its sign carried
__host__ __device__ inline uint64_t tcompose(uint64_t A, uint64_t B, int n) {
  const uint32_t nm = (1u << n) - 1;
  uint64_t C = 0;
  for (int i = 0; i < 2 * n; i++) {
    const uint32_t w = trow(A, n, i);
    C = tset(C, n, i, timage(B, n, w & nm, (w >> n) & nm) ^ (((w >> (2 * n)) & 1) << (2 * n)));
  }
  return C;
}

// the inverse: its symplectic part is Omega S^T Omega (S the tableau's bit matrix), and the sign of row i is the one
// with which the Clifford maps that row's Pauli back onto generator i
__host__ __device__ inline uint64_t tinverse(uint64_t C, int n) {
  const int m = 2 * n;
  const uint32_t nm = (1u << n) - 1;
  uint64_t D = 0;
  for (int i = 0; i < m; i++) {
    const int wi = (i + n) % m;
    uint32_t v = 0;
    for (int j = 0; j < m; j++) {
      const uint32_t sr = trow(C, n, (j + n) % m);
      v |= (((sr & nm) | (((sr >> n) & nm) << n)) >> wi & 1) << j;
    }
    const uint32_t x = v & nm, z = v >> n;
    D = tset(D, n, i, x | (z << n) | (((timage(C, n, x, z) >> (2 * n)) & 1) << (2 * n)));
  }
  return D;
}

// distance of one tableau (scalar probe over the bucket layout), -1 when absent
__device__ inline int find1(Table T, uint64_t key) {
  uint64_t b = hmix(key) & T.bmask;
  for (uint64_t guard = 0; guard <= T.bmask; guard++) {
    for (int j = 0; j < 8; j++) {
      const unsigned long long v = T.slot[b * 8 + j];
      if (!v) return -1;
      if ((v >> 8) == key) return (int)(v & 255);
    }
    b = (b + 1) & T.bmask;
  }
  return -1;
}

struct Cand { uint64_t K, Kinv; int qp, neg; };

// element i (sorted order) is a placement for a T on qubit q: flag = q' + 1 where K Z_q' K^-1 = +-Z_q, else 0
__global__ void k_cand_flags(const uint64_t *sorted, long long order, int n, int q, uint8_t *flag) {
  const uint32_t nm = (1u << n) - 1;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < order; i += (long long)gridDim.x * blockDim.x) {
    const uint64_t K = sorted[i];
    uint8_t f = 0;
    for (int qp = 0; qp < n; qp++) {
      const uint32_t w = trow(K, n, n + qp);
      if ((w & nm) == 0 && ((w >> n) & nm) == (1u << q)) f = (uint8_t)(qp + 1);
    }
    flag[i] = f;
  }
}
struct FlagNZ {
  const uint8_t *f;
  __device__ bool operator()(long long i) const { return f[i] != 0; }
};
__global__ void k_cand_build(const uint64_t *sorted, const uint32_t *idx, long long nc, const uint8_t *flag, int n,
                             Cand *out) {
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (long long)gridDim.x * blockDim.x) {
    const uint64_t K = sorted[idx[c]];
    const int qp = flag[idx[c]] - 1;
    const uint32_t w = trow(K, n, n + qp);
    out[c] = Cand{K, tinverse(K, n), qp, (int)((w >> (2 * n)) & 1)};
  }
}

// T = Tdg S: a flipped placement's inverse takes an S (Sdg when the placed T is a dagger) after it
__host__ __device__ inline uint64_t flipped_inv(const Cand &c, int dagger, int n) {
  return apply(c.Kinv, n, dagger ? 2 : 1, c.qp, 0);
}

constexpr int TOPK = 256;   // largest beam

// the block's B smallest values among best[0..256) and vals[0..256): bitonic sort of the 512 in LDS
__device__ inline void merge_topk(unsigned long long *best) {
  const int tid = threadIdx.x;
  for (int size = 2; size <= 512; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (int e = tid; e < 256; e += blockDim.x) {   // 256 compare-exchanges per stage
        const int lo = (e / stride) * stride * 2 + (e % stride), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const unsigned long long a = best[lo], b = best[hi];
        if ((a > b) == up) { best[lo] = b; best[hi] = a; }
      }
    }
  __syncthreads();
}

struct BeamArgs {
  Table T;
  int n;
  const uint64_t *Kprev;    // placements kept by the previous step (nb of them)
  const int *bcost;         // their summed segment lengths
  int nb, first;
  uint64_t D;               // this T gate's preceding segment
  const Cand *C;
  long long nc;
  int dg;                   // this T gate's dagger before any flip
  int B;
  unsigned long long *out;  // per block: its B smallest (cost << 40 | state << 32 | candidate << 1 | flip)
};

// one beam step (src/clifford.cu:190-201): every (state, placement, flip) costed by the table, each block keeping
// the B smallest of its contiguous range of items in LDS
__global__ __launch_bounds__(256) void k_beam(BeamArgs a) {
  __shared__ unsigned long long best[512];
  const int tid = threadIdx.x;
  best[tid] = ~0ull;
  best[256 + tid] = ~0ull;
  __syncthreads();
  const long long tot = (long long)a.nb * a.nc * 2;
  const long long per = (tot + gridDim.x - 1) / gridDim.x;
  const long long lo = (long long)blockIdx.x * per, hi = lo + per < tot ? lo + per : tot;
  for (long long base = lo; base < hi; base += 256) {
    const long long i = base + tid;
    unsigned long long v = ~0ull;
    if (i < hi) {
      const int b = (int)(i / (a.nc * 2));
      const long long c2 = i % (a.nc * 2), c = c2 >> 1;
      const uint64_t left = a.first ? a.D : tcompose(a.Kprev[b], a.D, a.n);
      const Cand cd = a.C[c];
      const uint64_t kinv = (c2 & 1) ? flipped_inv(cd, a.dg ^ cd.neg, a.n) : cd.Kinv;
      const int d = find1(a.T, tcompose(left, kinv, a.n));
      const unsigned long long cost = (unsigned long long)(a.bcost[b] + (d < 0 ? 999 : d));
      v = (cost << 40) | ((unsigned long long)b << 32) | (unsigned long long)c2;
    }
    // merge only when some new value beats the block's current B-th smallest
    if (__syncthreads_or(v < best[a.B - 1])) {
      best[256 + tid] = v;
      merge_topk(best);
    }
  }
  __syncthreads();
  for (int j = tid; j < a.B; j += blockDim.x) a.out[(long long)blockIdx.x * a.B + j] = best[j];
}

// the last segment after each kept state: its optimal length
__global__ void k_beam_final(Table T, int n, const uint64_t *Kprev, int nb, uint64_t D, int *out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) { const int d = find1(T, tcompose(Kprev[b], D, n)); out[b] = d < 0 ? 999 : d; }
}
