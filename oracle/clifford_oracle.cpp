// ORACLE (test infrastructure only): CPU restatement of the reference's Clifford-group table, the parity
// checker of the HIP Clifford table.  Only tests/ and bench.py's cpu_baseline leg may load it.
//
// Followed reference semantics (file:line in the reference tree):
//   stabiliser tableau, gate conjugation rules ..... src/clifford.cu:41-86 (own packing, same algebra)
//   gate set (H, S, Sdg, X, Z per qubit; CX pairs) .. src/clifford.cu:147-155
//   BFS by gate count from the identity ............. src/clifford.cu:251-280
//   optimal word read-back .......................... src/clifford.cu:354-372
//   cliffopt: placements, beam, final choice ........ src/clifford.cu:168-206, :285-385 (placements taken in
//                                                     sorted tableau order, the order the HIP path uses)
// Parity pins: the group orders 24 / 11,520 / 92,897,280 (closed form 2^(n^2+2n) prod (4^j - 1),
// src/clifford.cu:252), brute-force products of words, and gate identities (H^2 = S^4 = I, S Sdg = I).
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

namespace oc {

// A tableau packs 2n rows (images of X_0..X_{n-1}, then Z_0..Z_{n-1}); row = x bits | z bits << n | sign << 2n.
struct Tab {
  int n;
  int rb() const { return 2 * n + 1; }
  uint32_t row(uint64_t T, int i) const { return (uint32_t)((T >> (i * rb())) & ((1ull << rb()) - 1)); }
  uint64_t put(uint64_t T, int i, uint32_t r) const {
    const uint64_t m = ((1ull << rb()) - 1) << (i * rb());
    return (T & ~m) | ((uint64_t)r << (i * rb()));
  }
  uint64_t identity() const {
    uint64_t T = 0;
    for (int i = 0; i < n; i++) { T = put(T, i, 1u << i); T = put(T, n + i, 1u << (n + i)); }
    return T;
  }
  // G P G^dag of one row for gate kind k on qubit a (target b for CX)
  uint32_t conj(uint32_t w, int k, int a, int b) const {
    const uint32_t nm = (1u << n) - 1;
    uint32_t x = w & nm, z = (w >> n) & nm, r = (w >> (2 * n)) & 1;
    const uint32_t xa = (x >> a) & 1, za = (z >> a) & 1;
    switch (k) {
      case 0: r ^= xa & za; x = (x & ~(1u << a)) | (za << a); z = (z & ~(1u << a)) | (xa << a); break;  // H
      case 1: r ^= xa & za; z ^= xa << a; break;                                                      // S
      case 2: r ^= xa & (za ^ 1); z ^= xa << a; break;                                                // Sdg
      case 3: r ^= za; break;                                                                         // X
      case 4: r ^= xa; break;                                                                         // Z
      case 5: {                                                                                       // CX a->b
        const uint32_t xb = (x >> b) & 1, zb = (z >> b) & 1;
        r ^= xa & zb & (xb ^ za ^ 1);
        x ^= xa << b;
        z ^= zb << a;
        break;
      }
      case 6: r ^= xa ^ za; break;                                                                    // Y
    }
    return x | (z << n) | (r << (2 * n));
  }
  uint64_t apply(uint64_t T, int k, int a, int b) const {
    for (int i = 0; i < 2 * n; i++) T = put(T, i, conj(row(T, i), k, a, b));
    return T;
  }
};

// image of the Hermitian Pauli i^(x.z) X^x Z^z: product of the rows of its factors, X factors first, the phase
// counted in powers of i ((X^a Z^b)(X^c Z^d) = (-1)^(b.c) X^(a^c) Z^(b^d))
static uint32_t image(const Tab &tb, uint64_t T, uint32_t x, uint32_t z) {
  const int n = tb.n;
  const uint32_t nm = (1u << n) - 1;
  int k = __builtin_popcount(x & z);
  uint32_t ax = 0, az = 0;
  for (int q = 0; q < n; q++)
    if ((x >> q) & 1) {
      const uint32_t w = tb.row(T, q), bx = w & nm, bz = (w >> n) & nm;
      k += 2 * (int)((w >> (2 * n)) & 1) + __builtin_popcount(bx & bz) + 2 * __builtin_popcount(az & bx);
      ax ^= bx; az ^= bz;
    }
  for (int q = 0; q < n; q++)
    if ((z >> q) & 1) {
      const uint32_t w = tb.row(T, n + q), bx = w & nm, bz = (w >> n) & nm;
      k += 2 * (int)((w >> (2 * n)) & 1) + __builtin_popcount(bx & bz) + 2 * __builtin_popcount(az & bx);
      ax ^= bx; az ^= bz;
    }
  return ax | (az << n) | ((uint32_t)(((k - __builtin_popcount(ax & az)) & 3) >> 1) << (2 * n));
}
// A then B
static uint64_t compose(const Tab &tb, uint64_t A, uint64_t B) {
  const int n = tb.n;
  const uint32_t nm = (1u << n) - 1;
  uint64_t C = 0;
  for (int i = 0; i < 2 * n; i++) {
    const uint32_t w = tb.row(A, i);
    C = tb.put(C, i, image(tb, B, w & nm, (w >> n) & nm) ^ (w & (1u << (2 * n))));
  }
  return C;
}
// the inverse by brute force over the rows: row i of C^-1 is the Pauli h (with its sign) that C maps to the
// generator g_i, found among the 4^n Hermitian Paulis
static uint64_t inverse(const Tab &tb, uint64_t C) {
  const int n = tb.n;
  uint64_t D = 0;
  for (int i = 0; i < 2 * n; i++) {
    const uint32_t gx = i < n ? 1u << i : 0, gz = i < n ? 0 : 1u << (i - n);
    for (uint32_t v = 1; v < (1u << (2 * n)); v++) {
      const uint32_t x = v & ((1u << n) - 1), z = v >> n;
      const uint32_t im = image(tb, C, x, z);
      if ((im & ((1u << (2 * n)) - 1)) == (gx | (gz << n))) { D = tb.put(D, i, x | (z << n) | (im & (1u << (2 * n)))); break; }
    }
  }
  return D;
}

struct GateSet { int ng = 0; int k[32], a[32], b[32]; };
static GateSet gate_set(int n) {
  GateSet G;
  for (int q = 0; q < n; q++)
    for (int k : {0, 1, 2, 3, 4}) { G.k[G.ng] = k; G.a[G.ng] = q; G.b[G.ng] = 0; G.ng++; }
  for (int a = 0; a < n; a++)
    for (int b = 0; b < n; b++)
      if (a != b) { G.k[G.ng] = 5; G.a[G.ng] = a; G.b[G.ng] = b; G.ng++; }
  return G;
}
static int inv_kind(int k) { return k == 1 ? 2 : k == 2 ? 1 : k; }

static inline uint64_t mix(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
  return x;
}

// key -> distance, open addressing with linear probing; slot word = key << 8 | distance, 0 = empty
struct Table {
  std::vector<std::atomic<uint64_t>> slot;
  uint64_t mask = 0;
  explicit Table(int lg) : slot(1ull << lg), mask((1ull << lg) - 1) {
    for (auto &s : slot) s.store(0, std::memory_order_relaxed);
  }
  int find(uint64_t key) const {
    for (uint64_t i = mix(key) & mask;; i = (i + 1) & mask) {
      const uint64_t v = slot[i].load(std::memory_order_relaxed);
      if (!v) return -1;
      if ((v >> 8) == key) return (int)(v & 255);
    }
  }
  bool insert(uint64_t key, int d) {   // true if new
    const uint64_t w = (key << 8) | (uint64_t)d;
    for (uint64_t i = mix(key) & mask;; i = (i + 1) & mask) {
      uint64_t v = slot[i].load(std::memory_order_relaxed);
      if (!v) {
        uint64_t z = 0;
        if (slot[i].compare_exchange_strong(z, w)) return true;
        v = z;
      }
      if ((v >> 8) == key) return false;
    }
  }
};

struct Cliff {
  int n;
  Tab tb;
  GateSet G;
  Table table;
  std::vector<long> levels;
  Cliff(int n_, int lg) : n(n_), tb{n_}, G(gate_set(n_)), table(lg) {}
};

}  // namespace oc

extern "C" {

// BFS of the whole n-qubit Clifford group (n <= 3).  Returns the context; levels via orc_cliff_levels.
void *orc_cliff_build(int n, int threads) {
  if (n < 1 || n > 3) return nullptr;
  if (threads > 0) omp_set_num_threads(threads);
  const long order = n == 1 ? 24 : n == 2 ? 11520 : 92897280;
  int lg = 4;
  while ((1L << lg) < order * 4 / 3) lg++;
  oc::Cliff *c = new oc::Cliff(n, lg);
  const uint64_t id = c->tb.identity();
  c->table.insert(id, 0);
  std::vector<uint64_t> front{id};
  c->levels.push_back(1);
  for (int d = 1; !front.empty(); d++) {
    std::vector<std::vector<uint64_t>> part(omp_get_max_threads());
#pragma omp parallel
    {
      std::vector<uint64_t> &mine = part[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 4096)
      for (long i = 0; i < (long)front.size(); i++)
        for (int g = 0; g < c->G.ng; g++) {
          const uint64_t t = c->tb.apply(front[i], c->G.k[g], c->G.a[g], c->G.b[g]);
          if (c->table.insert(t, d)) mine.push_back(t);
        }
    }
    front.clear();
    for (auto &p : part) front.insert(front.end(), p.begin(), p.end());
    if (!front.empty()) c->levels.push_back((long)front.size());
  }
  return c;
}

int orc_cliff_levels(void *h, long *out, int maxd) {
  oc::Cliff *c = (oc::Cliff *)h;
  const int nl = (int)c->levels.size();
  for (int i = 0; i < nl && i < maxd; i++) out[i] = c->levels[i];
  return nl;
}

// order-independent digest of the table: sum over entries of mix(key * 64 + distance)
uint64_t orc_cliff_digest(void *h) {
  oc::Cliff *c = (oc::Cliff *)h;
  uint64_t s = 0;
#pragma omp parallel for reduction(+ : s)
  for (long i = 0; i < (long)c->table.slot.size(); i++) {
    const uint64_t v = c->table.slot[i].load(std::memory_order_relaxed);
    if (v) s += oc::mix((v >> 8) * 64 + (v & 255));
  }
  return s;
}

// tableau of a word of gate-set indices applied in order (gate word[0] first)
uint64_t orc_cliff_word(void *h, const int *word, int len) {
  oc::Cliff *c = (oc::Cliff *)h;
  uint64_t T = c->tb.identity();
  for (int i = 0; i < len; i++) T = c->tb.apply(T, c->G.k[word[i]], c->G.a[word[i]], c->G.b[word[i]]);
  return T;
}

int orc_cliff_find(void *h, uint64_t key) { return ((oc::Cliff *)h)->table.find(key); }

// optimal word for tableau T: step to a predecessor one gate closer, first in gate order
// (src/clifford.cu:354-372); returns its length or -1
int orc_cliff_readback(void *h, uint64_t T, int *out, int cap) {
  oc::Cliff *c = (oc::Cliff *)h;
  int d = c->table.find(T);
  if (d < 0 || d > cap) return -1;
  const int len = d;
  while (d > 0) {
    int g = 0;
    uint64_t pre = 0;
    for (; g < c->G.ng; g++) {
      pre = c->tb.apply(T, oc::inv_kind(c->G.k[g]), c->G.a[g], c->G.b[g]);
      if (c->table.find(pre) == d - 1) break;
    }
    if (g == c->G.ng) return -1;
    out[d - 1] = g;
    T = pre;
    d--;
  }
  return len;
}

int orc_cliff_ngates(void *h) { return ((oc::Cliff *)h)->G.ng; }

// cliffopt (src/clifford.cu:285-385) on the job text: the reference's stdout without its "seconds" field
// ({"clifford_gates": .., "clifford_gates_before": .., "t_gates": ..}) and the gate lines; placements are taken in
// sorted tableau order.  Returns 0, 2 bad job, 3 qubit count, 4 out too small or reconstruction failure.
int orc_cliff_opt(void *h, const char *job, long len, int beam, char *out, long cap) {
  oc::Cliff *c = (oc::Cliff *)h;
  const oc::Tab &tb = c->tb;
  const int n = c->n;
  std::istringstream in(std::string(job, (size_t)len));
  int jn = 0, t = 0;
  if (!(in >> jn >> t) || t < 0) return 2;
  if (jn != n) return 3;
  std::vector<int> tq(t), td(t);
  for (int i = 0; i < t; i++) if (!(in >> tq[i] >> td[i])) return 2;
  std::vector<uint64_t> D(t + 1);
  std::vector<int> seglen(t + 1);
  static const char *names[] = {"h", "s", "sdg", "x", "z", "cx", "y"};
  for (int j = 0; j <= t; j++) {
    int m = 0;
    if (!(in >> m)) return 2;
    seglen[j] = m;
    uint64_t X = tb.identity();
    for (int i = 0; i < m; i++) {
      std::string nm;
      int a = 0, b = 0, k = -1;
      if (!(in >> nm >> a)) return 2;
      for (int q = 0; q < 7; q++) if (nm == names[q]) k = q;
      if (k < 0) return 2;
      if (k == 5 && !(in >> b)) return 2;
      X = tb.apply(X, k, a, b);
    }
    D[j] = X;
  }
  beam = std::max(1, std::min(255, beam));
  // every element, sorted; placements of a T on qubit q: K with K Z_q' K^-1 = +-Z_q
  std::vector<uint64_t> elems;
  for (auto &sl : c->table.slot) { const uint64_t v = sl.load(std::memory_order_relaxed); if (v) elems.push_back(v >> 8); }
  std::sort(elems.begin(), elems.end());
  struct Cand { uint64_t K, Kinv; int qp, neg; };
  std::vector<std::vector<Cand>> cands(n);
  const uint32_t nm = (1u << n) - 1;
  for (uint64_t K : elems)
    for (int qp = 0; qp < n; qp++) {
      const uint32_t w = tb.row(K, n + qp);
      if ((w & nm) == 0 && ((w >> n) & nm) != 0 && __builtin_popcount((w >> n) & nm) == 1) {
        const int q = __builtin_ctz((w >> n) & nm);
        cands[q].push_back(Cand{K, oc::inverse(tb, K), qp, (int)((w >> (2 * n)) & 1)});
      }
    }
  auto flipped = [&](const Cand &cd, int dagger) { return tb.apply(cd.Kinv, dagger ? 2 : 1, cd.qp, 0); };
  std::vector<std::vector<int>> parent(t), flips(t);
  std::vector<std::vector<Cand>> chosen(t);
  std::vector<uint64_t> Kprev(1, 0);
  std::vector<int> bcost(1, 0);
  for (int j = 0; j < t; j++) {
    const std::vector<Cand> &C = cands[tq[j]];
    const long nc = (long)C.size(), nb = (long)Kprev.size(), tot = nb * nc * 2;
    std::vector<uint64_t> packed(tot);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < tot; i++) {
      const long b = i / (nc * 2), c2 = i % (nc * 2);
      const Cand &cd = C[c2 >> 1];
      const uint64_t left = j == 0 ? D[j] : oc::compose(tb, Kprev[b], D[j]);
      const uint64_t kinv = (c2 & 1) ? flipped(cd, td[j] ^ cd.neg) : cd.Kinv;
      const int d = c->table.find(oc::compose(tb, left, kinv));
      packed[i] = ((uint64_t)(bcost[b] + (d < 0 ? 999 : d)) << 40) | ((uint64_t)b << 32) | (uint64_t)c2;
    }
    const long keep = std::min<long>(beam, tot);
    std::partial_sort(packed.begin(), packed.begin() + keep, packed.end());
    parent[j].resize(keep); flips[j].resize(keep); chosen[j].resize(keep);
    std::vector<uint64_t> nk(keep);
    std::vector<int> nc2(keep);
    for (long i = 0; i < keep; i++) {
      const int b = (int)((packed[i] >> 32) & 0xff);
      const long c2 = (long)(packed[i] & 0xffffffffull);
      chosen[j][i] = C[c2 >> 1];
      parent[j][i] = b;
      flips[j][i] = (int)(c2 & 1);
      nk[i] = C[c2 >> 1].K;
      nc2[i] = (int)(packed[i] >> 40);
    }
    Kprev = nk;
    bcost = nc2;
  }
  int best = 0, bestv = 1 << 30;
  if (t > 0)
    for (size_t b = 0; b < Kprev.size(); b++) {
      const int d = c->table.find(oc::compose(tb, Kprev[b], D[t]));
      const int v = bcost[b] + (d < 0 ? 999 : d);
      if (v < bestv) { bestv = v; best = (int)b; }
    }
  std::vector<Cand> K(t);
  std::vector<int> flip(t), dag(t);
  for (int j = t - 1, b = best; j >= 0; j--) { K[j] = chosen[j][b]; flip[j] = flips[j][b]; b = parent[j][b]; }
  for (int j = 0; j < t; j++) {
    const int d0 = td[j] ^ K[j].neg;
    if (flip[j]) K[j].Kinv = flipped(K[j], d0);
    dag[j] = d0 ^ flip[j];
  }
  std::string s;
  int total = 0, orig = 0;
  std::vector<std::vector<int>> words(t + 1);
  for (int j = 0; j <= t; j++) {
    uint64_t X = j > 0 ? oc::compose(tb, K[j - 1].K, D[j]) : D[j];
    if (j < t) X = oc::compose(tb, X, K[j].Kinv);
    std::vector<int> w(64);
    const int L = orc_cliff_readback(h, X, w.data(), 64);
    if (L < 0) return 4;
    w.resize(L);
    words[j] = w;
    total += L;
    orig += seglen[j];
  }
  char buf[160];
  std::snprintf(buf, sizeof buf, "{\"clifford_gates\": %d, \"clifford_gates_before\": %d, \"t_gates\": %d}\n", total, orig, t);
  s = buf;
  for (int j = 0; j <= t; j++) {
    for (int g : words[j]) {
      if (c->G.k[g] == 5) std::snprintf(buf, sizeof buf, "cx %d %d\n", c->G.a[g], c->G.b[g]);
      else std::snprintf(buf, sizeof buf, "%s %d\n", names[c->G.k[g]], c->G.a[g]);
      s += buf;
    }
    if (j < t) { std::snprintf(buf, sizeof buf, "%s %d\n", dag[j] ? "tdg" : "t", K[j].qp); s += buf; }
  }
  if ((long)s.size() + 1 > cap) return 4;
  std::memcpy(out, s.data(), s.size());
  out[s.size()] = 0;
  return 0;
}

void orc_cliff_free(void *h) { delete (oc::Cliff *)h; }

}  // extern "C"
