// Kernels and table storage of the gate-count search modes beyond the plain level search: the distance table of the last K gates (`--table K`,
// src/cheese.cu:1431-1593, dispatch :1804-1842) and the meet-in-the-middle join (`--mitm`, :1595-1760).
// Included by search.hip inside namespace chz, before the level step (whose table build hooks use them); the
// drivers are in table_join.h.
//
// Distance table.  Every symmetry class of unitaries within K gates, at the level of its exact gate count: a
// search from the identity without bounds whose levels drop the classes already seen at a lower level (hash
// probe, then an exact class comparison).  Every level stays resident in HBM (residual, exponent, class key; 1 KB
// per class at n = 3: K = 8 is 17.9 M classes, about 19 GB), with an open-addressing multimap class key -> (level,
// node) over all of them, so the filter and the search's lookups compare residuals directly instead of replaying
// paths.  A search with the table runs its levels 1..D-K as usual; after each level t every frontier node is
// looked up, and a node whose class sits in the table at a level c <= D - t gives a circuit of t + c gates (the
// smallest (c, node) first; the continuation is the table entry's sequence reversed, inverted and mapped by the
// first group element relating the two residuals; the circuit is confirmed exactly before it is reported).
//
// Meet in the middle.  A forward search on U and a backward search on U^dag, symmetry classes off, each pruned
// against the full budget D; the side with the smaller frontier grows until the depths add up to D.  Then every
// backward node q (levels 1..b, residual R' = U^dag q^-1) is joined with the forward level a on U R' = R: a
// multimap of the forward class keys, probed with the key of U R' and confirmed by exact comparison on the
// device; the first backward node in level order with a match gives the circuit p followed by q^-1.

// ---------------------------------------------------------------------------------------------------------------
// device side

struct GTDev {                   // the table as the kernels see it
  const uint64_t *hkey;          // slot: class key | 1 (0: empty)
  const uint64_t *hval;          // slot: level << 32 | node
  uint64_t mask;                 // slots - 1
  const int4 *const *res;        // per level: residuals, E entries per node
  const int *const *pk;          // per level: sqrt2 exponents
};

__device__ __forceinline__ uint64_t gt_home(uint64_t key, uint64_t mask) { return mix64(key ^ 0x5bd1e9955bd1e995ull) & mask; }

// class keys of m residuals (the expansion's key for instance slot 0): one lane group per node
template <int NQ> __global__ __launch_bounds__(256) void k_node_keys(KSpec sp, const int4 *res, const int *pk,
                                                                    long long m, uint64_t *keys) {
  constexpr int E = 1 << (2 * NQ), GS = E, CPW = 64 / GS;
  const int lane = threadIdx.x & 63, le = lane & (GS - 1);
  const long long wave = wave_index();
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long b = wave * CPW; b < m; b += nwaves * CPW) {       // group-uniform trip count
    const long long j = b + lane / GS;
    const bool in = j < m;
    const int4 s = in ? res[j * E + le] : make_int4(0, 0, 0, 0);
    const uint64_t key = class_key<NQ>(s, in ? pk[j] : 0, le, sp);
    if (in && le == 0) keys[j] = key;
  }
}

// m keys of one level into the multimap: linear probing from the key's home slot, first empty slot claimed
__global__ void k_gt_insert(const uint64_t *keys, long long m, uint32_t level, uint64_t *hkey, uint64_t *hval,
                            uint64_t mask) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (long long)gridDim.x * blockDim.x) {
    const unsigned long long stored = keys[j] | 1ull;
    uint64_t pos = gt_home(keys[j], mask);
    for (uint64_t step = 0; step <= mask; step++, pos = (pos + 1) & mask) {
      if (atomicCAS((unsigned long long *)&hkey[pos], 0ull, stored) == 0ull) {
        hval[pos] = ((uint64_t)level << 32) | (uint64_t)j;
        break;
      }
    }
  }
}

// the table build's cross-level filter: a kept survivor whose class the table already holds (a lower level) is
// dropped (src/cheese.cu:1084-1107, SeenBefore).  One lane group per survivor; the child is rebuilt from its parent.
template <int NQ> __global__ __launch_bounds__(256) void k_gt_filter(KSpec sp, GTDev t, const uint64_t *surv_o,
                                                                    const uint64_t *surv_key, long long n,
                                                                    const int4 *res, const int *pk, uint8_t *kept) {
  constexpr int E = 1 << (2 * NQ), GS = E, CPW = 64 / GS;
  const int lane = threadIdx.x & 63, le = lane & (GS - 1);
  const long long wave = wave_index();
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  const uint64_t G = (uint64_t)sp.G;
  for (long long b = wave * CPW; b < n; b += nwaves * CPW) {
    const long long i = b + lane / GS;
    if (i >= n || !kept[i]) continue;                 // uniform within the lane group
    const uint64_t o = surv_o[i], key = surv_key[i];
    const long long pi = (long long)(o / G);
    int4 s;
    int k;
    make_child<NQ>(res + pi * E, pk[pi], sp.gates[(int)(o % G)], le, s, k);
    const uint64_t stored = key | 1ull;
    uint64_t pos = gt_home(key, t.mask);
    bool dup = false;
    for (uint64_t step = 0; step <= t.mask && !dup; step++, pos = (pos + 1) & t.mask) {
      const uint64_t h = t.hkey[pos];
      if (h == 0) break;
      if (h != stored) continue;
      const uint64_t v = t.hval[pos];
      const int l = (int)(v >> 32);
      const long long j = (long long)(v & 0xffffffffull);
      dup = same_class<NQ>(s, k, t.res[l][j * E + le], t.pk[l][j], le, sp);
    }
    if (dup && le == 0) kept[i] = 0;
  }
}

// the search's lookups after level t: node j of the frontier whose class is in the table at a level c <= maxc
// gets cost[j] = c and hit[j] = the table entry (level << 32 | node); best = min over nodes of c << 40 | j
template <int NQ> __global__ __launch_bounds__(256) void k_gt_probe(KSpec sp, GTDev t, const int4 *res, const int *pk,
                                                                   const uint64_t *keys, long long m, int maxc,
                                                                   uint8_t *cost, uint64_t *hit,
                                                                   unsigned long long *best) {
  constexpr int E = 1 << (2 * NQ), GS = E, CPW = 64 / GS;
  const int lane = threadIdx.x & 63, le = lane & (GS - 1);
  const long long wave = wave_index();
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long b = wave * CPW; b < m; b += nwaves * CPW) {
    const long long j = b + lane / GS;
    if (j >= m) continue;
    const int4 s = res[j * E + le];
    const int k = pk[j];
    const uint64_t stored = keys[j] | 1ull;
    uint64_t pos = gt_home(keys[j], t.mask);
    int c = 255;
    uint64_t at = 0;
    for (uint64_t step = 0; step <= t.mask; step++, pos = (pos + 1) & t.mask) {
      const uint64_t h = t.hkey[pos];
      if (h == 0) break;
      if (h != stored) continue;
      const uint64_t v = t.hval[pos];
      const int l = (int)(v >> 32);
      if (l > maxc || l >= c) continue;
      const long long x = (long long)(v & 0xffffffffull);
      if (same_class<NQ>(t.res[l][x * E + le], t.pk[l][x], s, k, le, sp)) { c = l; at = v; }
    }
    if (le == 0) {
      cost[j] = (uint8_t)c;
      hit[j] = at;
      if (c <= maxc) atomicMin(best, ((unsigned long long)c << 40) | (unsigned long long)j);
    }
  }
}

// K = U R (normalised), U the forward target: the backward node's residual seen from the forward side
template <int NQ> __device__ __forceinline__ bool mul_left(const int4 *U, int ku, const int4 *R, int kr, int le,
                                                           int4 &out, int &k) {
  constexpr int N = 1 << NQ, GS = N * N;
  const int r = le >> NQ, c = le & (N - 1);
  long long a = 0, b = 0, cc = 0, d = 0;
  for (int m = 0; m < N; m++) {
    const int4 u = U[r * N + m], v = R[m * N + c];
    a += (long long)u.x * v.x - (long long)u.y * v.w - (long long)u.z * v.z - (long long)u.w * v.y;
    b += (long long)u.x * v.y + (long long)u.y * v.x - (long long)u.z * v.w - (long long)u.w * v.z;
    cc += (long long)u.x * v.z + (long long)u.y * v.y + (long long)u.z * v.x - (long long)u.w * v.w;
    d += (long long)u.x * v.w + (long long)u.y * v.z + (long long)u.z * v.y + (long long)u.w * v.x;
  }
  const bool of = a != (int)a || b != (int)b || cc != (int)cc || d != (int)d;
  out = make_int4((int)a, (int)b, (int)cc, (int)d);
  k = ku + kr;
  if (gany<GS>(of)) return false;
  while (k > 0 && gall<GS>(zhalf(out))) { out = zdown(out); k--; }
  return true;
}

// the join for one backward level: node j (residual R'_j) -> U R'_j -> key -> the forward multimap; hashes[j] = 1
// when some forward key is equal, fwd[j] = the forward node whose residual equals U R'_j exactly (or -1)
template <int NQ> __global__ __launch_bounds__(256) void k_mitm_join(KSpec sp, GTDev f, const int4 *U, int ku,
                                                                    const int4 *res, const int *pk, long long m,
                                                                    uint8_t *hashes, long long *fwd,
                                                                    unsigned long long *ovf) {
  constexpr int E = 1 << (2 * NQ), GS = E, CPW = 64 / GS;
  const int lane = threadIdx.x & 63, le = lane & (GS - 1);
  const long long wave = wave_index();
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long b = wave * CPW; b < m; b += nwaves * CPW) {
    const long long j = b + lane / GS;
    if (j >= m) continue;
    int4 K;
    int kk;
    if (!mul_left<NQ>(U, ku, res + j * E, pk[j], le, K, kk)) {
      if (le == 0) { atomicAdd(ovf, 1ull); hashes[j] = 0; fwd[j] = -1; }
      continue;
    }
    const uint64_t key = class_key<NQ>(K, kk, le, sp);
    const uint64_t stored = key | 1ull;
    uint64_t pos = gt_home(key, f.mask);
    bool any = false;
    long long match = -1;
    for (uint64_t step = 0; step <= f.mask && match < 0; step++, pos = (pos + 1) & f.mask) {
      const uint64_t h = f.hkey[pos];
      if (h == 0) break;
      if (h != stored) continue;
      any = true;
      const long long x = (long long)(f.hval[pos] & 0xffffffffull);
      if (f.pk[0][x] == kk && gall<GS>(zeq(f.res[0][x * E + le], K))) match = x;
    }
    if (le == 0) { hashes[j] = any; fwd[j] = match; }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// host side

// candidate index of each candidate's inverse (t <-> tdg, s <-> sdg, the rest self-inverse), by name
// (src/cheese.cu:1643-1656); -1 where the gate set lacks it
static std::vector<int> inverse_candidates(const HostSpec &hs) {
  std::vector<int> inv(hs.G(), -1);
  for (int g = 0; g < hs.G(); g++) {
    const std::string &nm = hs.gates[g].name;
    const size_t a = nm.find('_'), b = nm.find('_', a + 1);
    if (a == std::string::npos) continue;
    std::string gate = nm.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1);
    if (gate == "t") gate = "tdg"; else if (gate == "tdg") gate = "t";
    else if (gate == "s") gate = "sdg"; else if (gate == "sdg") gate = "s";
    const std::string want = nm.substr(0, a + 1) + gate + (b == std::string::npos ? "" : nm.substr(b));
    for (int h = 0; h < hs.G(); h++) if (hs.gates[h].name == want) inv[g] = h;
  }
  return inv;
}

// an open-addressing multimap over levels of resident residuals (the distance table; the forward level of a join)
struct GTable {
  int K = -1;
  std::vector<long long> count;               // nodes per level
  std::vector<int4 *> res;
  std::vector<int *> pk;
  std::vector<uint64_t *> keys;
  std::vector<std::vector<uint32_t>> par;     // host copies of the tree: parent in the level below, gate
  std::vector<std::vector<uint8_t>> gate;
  uint64_t *hkey = nullptr, *hval = nullptr;
  unsigned long long nslots = 0;
  const int4 **d_res = nullptr;               // device arrays of the level pointers (CHZ_DMAX + 1 entries)
  const int **d_pk = nullptr;
  std::vector<chz_level> lines;
  double seconds = 0;
  bool owns_levels = true;                    // false: the levels are a context's frontier (a join's forward side)
  long long total() const { long long s = 0; for (long long v : count) s += v; return s; }
  GTDev dev() const { return GTDev{hkey, hval, nslots - 1, d_res, d_pk}; }
  void release() {
    if (owns_levels) {
      for (auto p : res) dev_free(p);
      for (auto p : pk) dev_free(p);
    }
    for (auto p : keys) dev_free(p);
    dev_free(hkey); dev_free(hval); dev_free((void *)d_res); dev_free((void *)d_pk);
    res.clear(); pk.clear(); keys.clear(); count.clear(); par.clear(); gate.clear();
    hkey = hval = nullptr; d_res = nullptr; d_pk = nullptr; nslots = 0;
  }
};

// (re)sizes the multimap for `need` entries at load <= 1/2 and inserts every level's keys
static void gt_rehash(GTable &T, long long need, hipStream_t s) {
  unsigned long long slots = 1ull << 12;
  while ((long long)(slots / 2) < need) slots <<= 1;
  if (slots != T.nslots) {
    HIPCHK(hipStreamSynchronize(s));
    dev_free(T.hkey);
    dev_free(T.hval);
    T.hkey = (uint64_t *)dev_alloc(slots * 8);
    T.hval = (uint64_t *)dev_alloc(slots * 8);
    T.nslots = slots;
  }
  HIPCHK(hipMemsetAsync(T.hkey, 0, slots * 8, s));
  for (size_t l = 0; l < T.keys.size(); l++)
    if (T.count[l])
      hipLaunchKernelGGL(k_gt_insert, dim3(grid_for(T.count[l] / 64 + 1)), dim3(256), 0, s, T.keys[l], T.count[l],
                         (uint32_t)l, T.hkey, T.hval, (uint64_t)(T.nslots - 1));
  HIPCHK(hipGetLastError());
}

// appends a level of n nodes (residuals and exponents on the device, copied when the table owns its levels; tree
// arrays tp/tg on the device, or null for a root level) and inserts its class keys
template <int NQ> static void gt_add_level(const KSpec &ks, GTable &T, const int4 *res, const int *pk,
                                           const uint32_t *tp, const uint8_t *tg, long long n, hipStream_t s) {
  constexpr int E = 1 << (2 * NQ), CPW = 64 / E;
  if (!T.d_res) {
    T.d_res = (const int4 **)dev_alloc((CHZ_DMAX + 1) * sizeof(void *));
    T.d_pk = (const int **)dev_alloc((CHZ_DMAX + 1) * sizeof(void *));
  }
  const size_t l = T.count.size();
  if (l > CHZ_DMAX) { set_err("table deeper than 48 levels"); throw DeviceError(); }
  int4 *r = (int4 *)res;
  int *p = (int *)pk;
  if (T.owns_levels) {
    r = (int4 *)dev_alloc((size_t)std::max<long long>(n, 1) * E * sizeof(int4));
    p = (int *)dev_alloc((size_t)std::max<long long>(n, 1) * 4);
    if (n) {
      HIPCHK(hipMemcpyAsync(r, res, (size_t)n * E * sizeof(int4), hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemcpyAsync(p, pk, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    }
  }
  uint64_t *k = (uint64_t *)dev_alloc((size_t)std::max<long long>(n, 1) * 8);
  if (n) hipLaunchKernelGGL(k_node_keys<NQ>, dim3(grid_for((n + CPW - 1) / CPW)), dim3(256), 0, s, ks, r, p, n, k);
  T.res.push_back(r);
  T.pk.push_back(p);
  T.keys.push_back(k);
  T.count.push_back(n);
  T.par.emplace_back();
  T.gate.emplace_back();
  if (tp && n) {
    T.par.back().resize(n);
    T.gate.back().resize(n);
    HIPCHK(hipMemcpyAsync(T.par.back().data(), tp, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(T.gate.back().data(), tg, (size_t)n, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipMemcpyAsync((void *)(T.d_res + l), &T.res.back(), sizeof(void *), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync((void *)(T.d_pk + l), &T.pk.back(), sizeof(void *), hipMemcpyHostToDevice, s));
  const long long tot = T.total();
  if ((long long)(T.nslots / 2) < tot) gt_rehash(T, tot, s);
  else if (n)
    hipLaunchKernelGGL(k_gt_insert, dim3(grid_for(n / 64 + 1)), dim3(256), 0, s, k, n, (uint32_t)l, T.hkey, T.hval,
                       (uint64_t)(T.nslots - 1));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));     // the host tree copies and the pointer uploads read host memory
}

