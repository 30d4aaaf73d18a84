// ORACLE (test infrastructure only): CPU restatement of the reference's T-count search, the parity checker
// of the HIP T-count path.  Only tests/ and bench.py's cpu_baseline leg may load it.
//
// Followed reference semantics (file:line in the reference tree):
//   .tspec reader ........................................ src/tcount.cu:433-450
//   peel S <- R(P_p)^-1 S on the channel, renormalise ..... src/tcount.cu:80-112
//   right-coset label, exact coset equality ............... src/tcount.cu:113-188 (own hash; sorted columns)
//   level growth, dedupe (min index per coset), table filter src/tcount.cu:327-395
//   table from I (levels <= K), search from U, result ..... src/tcount.cu:452-532
// Parity pins: the reference's own search logs and results (benchmarks/optimal/*-tcount.json), the published
// table sizes (n=3 K=4: 5,605,741; n=2 K=6: 1,728,736), and exact verification of every decomposition.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace ot {

struct TSpec {
  int n = 0, P = 0, e = 0;
  std::vector<int> x, y, tr, ts;
};

// channel matrix: entries (x + y sqrt2) / sqrt2^e, P x P, rows = a (Pauli of the output), cols = b
struct Mat {
  int e = 0;
  bool ovf = false;
  std::vector<int32_t> x, y;
};

static inline uint64_t mix(uint64_t h) {
  h ^= h >> 32; h *= 0xd6e8feb86659fd93ULL; h ^= h >> 32; h *= 0xd6e8feb86659fd93ULL; return h ^ (h >> 32);
}

struct Ctx {
  TSpec s;
  int P;
  std::vector<int32_t> bx, by;   // base of the current tree (identity for the table, U for the search)
  int be;
};

// S <- R(P_p)^-1 S: rows commuting with P_p gain a factor sqrt2, an anticommuting row a becomes
// row_a - s row_r (i P_p P_a = s P_r); then (x + y s) / s = y + (x/2) s while every x is even
static void peel(Mat &M, int p, const TSpec &sp) {
  const int P = sp.P;
  std::vector<int32_t> nx(P * P), ny(P * P);
  for (int a = 0; a < P; a++) {
    const int r = sp.tr[(p - 1) * P + a], s = sp.ts[(p - 1) * P + a];
    for (int c = 0; c < P; c++) {
      int64_t vx, vy;
      if (r < 0) { vx = 2 * (int64_t)M.y[a * P + c]; vy = M.x[a * P + c]; }
      else { vx = (int64_t)M.x[a * P + c] - s * (int64_t)M.x[r * P + c]; vy = (int64_t)M.y[a * P + c] - s * (int64_t)M.y[r * P + c]; }
      if (vx > INT32_MAX || vx < INT32_MIN || vy > INT32_MAX || vy < INT32_MIN) M.ovf = true;
      nx[a * P + c] = (int32_t)vx;
      ny[a * P + c] = (int32_t)vy;
    }
  }
  M.x.swap(nx);
  M.y.swap(ny);
  M.e++;
  while (M.e > 0) {
    bool even = true;
    for (int32_t v : M.x) if (v & 1) { even = false; break; }
    if (!even) break;
    for (int i = 0; i < P * P; i++) { const int32_t t = M.x[i]; M.x[i] = M.y[i]; M.y[i] = t / 2; }
    M.e--;
  }
}

// sign-normalised column c: the sign of its first nonzero entry made positive
static void column(const Mat &M, int P, int c, std::vector<int64_t> &out) {
  int s = 1;
  for (int r = 0; r < P; r++) {
    const int32_t x = M.x[r * P + c], y = M.y[r * P + c];
    if (x || y) { s = (x > 0 || (x == 0 && y > 0)) ? 1 : -1; break; }
  }
  out.resize(2 * P);
  for (int r = 0; r < P; r++) { out[2 * r] = s * (int64_t)M.x[r * P + c]; out[2 * r + 1] = s * (int64_t)M.y[r * P + c]; }
}

// coset key: exponent and the sum of column hashes (independent of column order and signs)
static uint64_t key(const Mat &M, int P) {
  uint64_t sum = 0;
  std::vector<int64_t> col;
  for (int c = 0; c < P; c++) {
    column(M, P, c, col);
    uint64_t h = 0x243f6a8885a308d3ULL;
    for (int64_t v : col) h = mix(h ^ (uint64_t)v) + 0x9e3779b97f4a7c15ULL;
    sum += h;
  }
  return mix(sum ^ ((uint64_t)(M.e + 1) * 0xc2b2ae3d27d4eb4fULL));
}

static bool same_coset(const Mat &A, const Mat &B, int P) {
  if (A.e != B.e) return false;
  std::vector<std::vector<int64_t>> ca(P), cb(P);
  for (int c = 0; c < P; c++) { column(A, P, c, ca[c]); column(B, P, c, cb[c]); }
  std::sort(ca.begin(), ca.end());
  std::sort(cb.begin(), cb.end());
  return ca == cb;
}

struct Node { uint32_t par; uint8_t p; };

struct Tree {
  std::vector<int32_t> bx, by;
  int be = 0;
  std::vector<std::vector<Node>> lv;   // lv[0] = {root}
  std::vector<std::vector<uint64_t>> keys;
  Mat build(const TSpec &sp, int t, uint32_t i) const {
    std::vector<int> seq(t);
    for (int l = t; l >= 1; l--) { seq[l - 1] = lv[l][i].p; i = lv[l][i].par; }
    Mat M;
    M.e = be; M.x = bx; M.y = by;
    for (int p : seq) peel(M, p, sp);
    return M;
  }
  std::vector<int> seq(int t, uint32_t i) const {
    std::vector<int> s(t);
    for (int l = t; l >= 1; l--) { s[l - 1] = lv[l][i].p; i = lv[l][i].par; }
    return s;
  }
};

struct TableIndex {                      // every table coset by key: (level, index)
  std::unordered_multimap<uint64_t, std::pair<int, uint32_t>> m;
};

static double now() { return omp_get_wtime(); }

// one more level of T; table: drop cosets already in idx (at a lower level).  Returns the kept count.
static long grow(const TSpec &sp, Tree &T, TableIndex *idx, bool table, std::string &out, const char *tag, int *ovf) {
  const double t0 = now();
  const int P = sp.P, G = P - 1, t = (int)T.lv.size(), tp = t - 1;
  const long m = (long)T.lv[tp].size(), nc = m * G;
  std::vector<uint64_t> ck(nc);
#pragma omp parallel for schedule(dynamic, 4)
  for (long i = 0; i < m; i++) {
    const Mat par = T.build(sp, tp, (uint32_t)i);
    for (int p = 1; p <= G; p++) {
      Mat c = par;
      peel(c, p, sp);
      if (c.ovf) *ovf = 1;
      ck[i * G + p - 1] = key(c, P);
    }
  }
  // exact dedupe: smallest child index of every coset
  std::vector<long> ord(nc);
  std::iota(ord.begin(), ord.end(), 0L);
  std::stable_sort(ord.begin(), ord.end(), [&](long a, long b) { return ck[a] < ck[b]; });
  std::vector<char> keep(nc, 1);
  std::vector<std::pair<long, long>> runs;
  for (long a = 0; a < nc;) {
    long b = a + 1;
    while (b < nc && ck[ord[b]] == ck[ord[a]]) b++;
    if (b - a > 1) runs.push_back({a, b});
    a = b;
  }
  auto child = [&](long o) { Mat c = T.build(sp, tp, (uint32_t)(o / G)); peel(c, (int)(o % G) + 1, sp); return c; };
#pragma omp parallel for schedule(dynamic, 1)
  for (long r = 0; r < (long)runs.size(); r++) {
    std::vector<Mat> reps;
    for (long j = runs[r].first; j < runs[r].second; j++) {
      const Mat c = child(ord[j]);
      bool dup = false;
      for (const Mat &q : reps) if (same_coset(c, q, P)) { dup = true; break; }
      if (dup) keep[ord[j]] = 0; else reps.push_back(c);
    }
  }
  std::vector<long> kept;
  for (long o = 0; o < nc; o++) if (keep[o]) kept.push_back(o);
  std::vector<char> fresh(kept.size(), 1);
  if (table && idx) {
#pragma omp parallel for schedule(dynamic, 64)
    for (long j = 0; j < (long)kept.size(); j++) {
      const uint64_t kk = ck[kept[j]];
      auto range = idx->m.equal_range(kk);
      if (range.first == range.second) continue;
      const Mat c = child(kept[j]);
      for (auto it = range.first; it != range.second; ++it)
        if (same_coset(c, T.build(sp, it->second.first, it->second.second), P)) { fresh[j] = 0; break; }
    }
  }
  std::vector<Node> lvl;
  std::vector<uint64_t> keys;
  for (long j = 0; j < (long)kept.size(); j++)
    if (fresh[j]) { lvl.push_back({(uint32_t)(kept[j] / G), (uint8_t)(kept[j] % G + 1)}); keys.push_back(ck[kept[j]]); }
  const long nf = (long)lvl.size();
  T.lv.push_back(std::move(lvl));
  T.keys.push_back(keys);
  if (table && idx)
    for (long j = 0; j < nf; j++) idx->m.insert({T.keys[t][j], {t, (uint32_t)j}});
  char buf[256];
  std::snprintf(buf, sizeof buf, "{%s\"t\": %d, \"parents\": %ld, \"children\": %ld, \"cosets\": %ld, \"new\": %ld, \"seconds\": %.2f}\n",
                tag, t, m, nc, (long)kept.size(), nf, now() - t0);
  out += buf;
  return nf;
}

static int read_spec(const std::string &text, TSpec &s) {
  std::istringstream in(text);
  std::string tok;
  in >> tok;
  if (tok != "tcount-spec") return 2;
  int ver; in >> ver;
  while (in >> tok) {
    if (tok == "n") { in >> s.n; s.P = 1 << (2 * s.n); }
    else if (tok == "channel") { in >> s.e; s.x.resize(s.P * s.P); s.y.resize(s.P * s.P); for (int i = 0; i < s.P * s.P; i++) in >> s.x[i] >> s.y[i]; }
    else if (tok == "rotations") { int m; in >> m; s.tr.resize(m * s.P); s.ts.resize(m * s.P); for (int i = 0; i < m * s.P; i++) in >> s.tr[i] >> s.ts[i]; }
    else if (tok == "end") break;
    else return 2;
  }
  if (s.n < 1 || s.n > 3) return 3;
  return 0;
}

struct State {
  TSpec sp;
  Tree table;
  TableIndex idx;
  int K = 0;
};

}  // namespace ot

extern "C" {

// Build the coset table (levels <= K from the identity) for the spec's n; the table lines go to out.
void *orc_tc_table(const char *spec, long len, int K, int threads, char *out, long cap) {
  if (threads > 0) omp_set_num_threads(threads);
  ot::State *st = new ot::State();
  if (ot::read_spec(std::string(spec, (size_t)len), st->sp)) { delete st; return nullptr; }
  const int P = st->sp.P;
  st->K = K;
  st->table.bx.assign(P * P, 0);
  st->table.by.assign(P * P, 0);
  for (int i = 0; i < P; i++) st->table.bx[i * P + i] = 1;
  st->table.be = 0;
  st->table.lv.assign(1, std::vector<ot::Node>{{0, 0}});
  ot::Mat I = st->table.build(st->sp, 0, 0);
  st->table.keys.assign(1, std::vector<uint64_t>{ot::key(I, P)});
  st->idx.m.insert({st->table.keys[0][0], {0, 0u}});
  std::string o;
  int ovf = 0;
  for (int l = 1; l <= K; l++)
    if (!ot::grow(st->sp, st->table, &st->idx, true, o, "\"dir\": \"table\", ", &ovf)) break;
  if (out && cap > 0) { const size_t n = std::min((size_t)cap - 1, o.size()); std::memcpy(out, o.data(), n); out[n] = 0; }
  return st;
}

long orc_tc_table_size(void *h) {
  ot::State *st = (ot::State *)h;
  long s = 0;
  for (auto &l : st->table.lv) s += (long)l.size();
  return s;
}

// Search the spec's target against the table (src/tcount.cu:497-531); the stdout lines go to out.
// Returns the reference exit code: 0 found (rot receives 2 * tcount ints: p, inv), 1 none <= maxt, 2/3 bad spec,
// 4 overflow.
int orc_tc_search(void *h, const char *spec, long len, int maxt, char *out, long cap, int *rot) {
  ot::State *st = (ot::State *)h;
  ot::TSpec sp;
  if (int rc = ot::read_spec(std::string(spec, (size_t)len), sp)) return rc;
  const int P = sp.P;
  const double t0 = ot::now();
  ot::Tree T;
  T.bx = sp.x; T.by = sp.y; T.be = sp.e;
  T.lv.assign(1, std::vector<ot::Node>{{0, 0}});
  ot::Mat U = T.build(sp, 0, 0);
  T.keys.assign(1, std::vector<uint64_t>{ot::key(U, P)});
  std::string o;
  int ovf = 0, rc = 1;
  const int Kt = (int)st->table.lv.size() - 1;
  for (int j = 0; j <= maxt; j++) {
    if (j > 0 && !ot::grow(sp, T, nullptr, false, o, "\"dir\": \"search\", ", &ovf)) break;
    if (ovf) { rc = 4; break; }
    const int maxc = std::min(Kt, maxt - j);
    const long n = (long)T.lv[j].size();
    std::vector<int> hitc(n, 1 << 30);
    std::vector<uint32_t> hiti(n, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < n; i++) {
      auto range = st->idx.m.equal_range(T.keys[j][i]);
      if (range.first == range.second) continue;
      const ot::Mat A = T.build(sp, j, (uint32_t)i);
      for (auto it = range.first; it != range.second; ++it) {
        const int lvl = it->second.first;
        if (lvl > maxc || lvl >= hitc[i]) continue;
        if (ot::same_coset(A, st->table.build(st->sp, lvl, it->second.second), P)) { hitc[i] = lvl; hiti[i] = it->second.second; }
      }
    }
    long best = -1;
    int bestc = 1 << 30;
    for (long i = 0; i < n; i++) if (hitc[i] < bestc) { bestc = hitc[i]; best = i; }
    if (best >= 0) {
      const std::vector<int> fw = T.seq(j, (uint32_t)best), tb = st->table.seq(bestc, hiti[best]);
      char buf[64];
      std::snprintf(buf, sizeof buf, "{\"tcount\": %d, \"rotations\": [", j + bestc);
      std::string s = buf;
      int k = 0;
      for (int p : fw) { s += (k ? ", [" : "[") + std::to_string(p) + ", 0]"; rot[2 * k] = p; rot[2 * k + 1] = 0; k++; }
      for (int i = (int)tb.size() - 1; i >= 0; i--) { s += (k ? ", [" : "[") + std::to_string(tb[i]) + ", 1]"; rot[2 * k] = tb[i]; rot[2 * k + 1] = 1; k++; }
      std::snprintf(buf, sizeof buf, "], \"split\": [%d, %d], \"seconds\": %.2f}\n", j, bestc, ot::now() - t0);
      o += s + buf;
      rc = 0;
      break;
    }
    if (j + Kt >= maxt) break;
  }
  if (rc == 1) {
    char buf[96];
    std::snprintf(buf, sizeof buf, "{\"tcount_above\": %d, \"seconds\": %.2f}\n", maxt, ot::now() - t0);
    o += buf;
  }
  if (out && cap > 0) { const size_t nn = std::min((size_t)cap - 1, o.size()); std::memcpy(out, o.data(), nn); out[nn] = 0; }
  return rc;
}

void orc_tc_free(void *h) { delete (ot::State *)h; }

}  // extern "C"
