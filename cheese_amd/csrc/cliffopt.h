// Host side of `cliffopt JOB [--beam B]` (src/clifford.cu:285-385): the job file, the placements of every T gate
// (in the table's sorted element order, so the search is deterministic), the beam over them, the final choice and
// the optimal gate lists of the new segments.  Included by clifford.hip inside its extern "C" block.

namespace {

struct CliffJob {
  int n = 0;
  std::vector<int> tq, td;                                   // T gates: qubit, dagger
  std::vector<std::vector<std::array<int, 3>>> seg;          // t + 1 segments of (kind, a, b)
};

// the job text (src/clifford.cu:224-238): "n t", t lines "q dagger", t + 1 lines "m gate q [q2] ..."
int read_cliff_job(const char *text, size_t len, CliffJob &J, std::string &err) {
  std::istringstream in(std::string(text, len));
  int t = -1;
  if (!(in >> J.n >> t) || J.n < 1 || J.n > 3 || t < 0 || t > 4096) { err = "bad job header"; return 2; }
  J.tq.resize(t);
  J.td.resize(t);
  for (int i = 0; i < t; i++)
    if (!(in >> J.tq[i] >> J.td[i]) || J.tq[i] < 0 || J.tq[i] >= J.n || (J.td[i] != 0 && J.td[i] != 1)) {
      err = "bad T gate line";
      return 2;
    }
  static const char *names[] = {"h", "s", "sdg", "x", "z", "cx", "y"};
  J.seg.resize(t + 1);
  for (auto &s : J.seg) {
    int m = -1;
    if (!(in >> m) || m < 0) { err = "bad segment length"; return 2; }
    for (int i = 0; i < m; i++) {
      std::string nm;
      if (!(in >> nm)) { err = "truncated segment"; return 2; }
      int k = -1;
      for (int j = 0; j < 7; j++) if (nm == names[j]) k = j;
      if (k < 0) { err = "unknown Clifford gate " + nm; return 2; }
      int a = 0, b = 0;
      if (!(in >> a) || a < 0 || a >= J.n) { err = "bad qubit"; return 2; }
      if (k == 5 && (!(in >> b) || b < 0 || b >= J.n || b == a)) { err = "bad cx target"; return 2; }
      s.push_back({k, a, b});
    }
  }
  return 0;
}

}  // namespace

int chz_cliff_opt(chz_cliff *c, const char *job, size_t len, int beam, char *out, size_t cap) {
  chz::RelaxedCapture relaxed;
  using chz_cl::Cand;
  const double t0 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  CliffJob J;
  std::string err;
  if (const int rc = read_cliff_job(job, len, J, err)) { chz_cl::g_err = err; return rc; }
  if (J.n != c->n) { chz_cl::g_err = "the job's qubit count differs from the table's"; return 3; }
  beam = std::max(1, std::min(chz_cl::TOPK - 1, beam));
  const int n = J.n, t = (int)J.tq.size();
  CLCHK(hipSetDevice(c->device));
  if (const int rc = cliff_sorted(c)) return rc;
  std::vector<uint64_t> D(t + 1);                             // segment tableaux
  for (int j = 0; j <= t; j++) {
    uint64_t X = chz_cl::identity(n);
    for (const auto &g : J.seg[j]) X = chz_cl::apply(X, n, g[0], g[1], g[2]);
    D[j] = X;
  }
  // placements per qubit, in sorted element order (ordered compaction of the flags)
  std::vector<Cand *> cands(n, nullptr);
  std::vector<long long> ncand(n, 0);
  std::vector<void *> owned;
  auto release = [&]() { for (void *p : owned) (void)cl_free(p); owned.clear(); };
  auto dalloc = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (cl_malloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
    owned.push_back(p);
    return p;
  };
  {
    const long long ord = c->order, nb = (ord + CT_TILE - 1) / CT_TILE;
    uint8_t *flag = (uint8_t *)dalloc(ord);
    unsigned long long *bc = (unsigned long long *)dalloc((nb + 1) * 8);
    unsigned long long *tmp = (unsigned long long *)dalloc((nb / SC_TILE + 8) * 8);
    uint32_t *idx = (uint32_t *)dalloc(ord * 4);
    if (!flag || !bc || !tmp || !idx) { release(); chz_cl::g_err = "device memory"; return 5; }
    for (int q = 0; q < n; q++) {
      if (std::find(J.tq.begin(), J.tq.end(), q) == J.tq.end()) continue;
      hipLaunchKernelGGL(chz_cl::k_cand_flags, dim3(4096), dim3(256), 0, c->stream, c->sorted, ord, n, q, flag);
      const chz_cl::FlagNZ f{flag};
      hipLaunchKernelGGL(k_flag_count<chz_cl::FlagNZ>, dim3((unsigned)nb), dim3(CT_THREADS), 0, c->stream, f, ord, bc);
      if (scan_excl(bc, nb, tmp, c->stream) != hipSuccess) { release(); chz_cl::g_err = "scan"; return 5; }
      hipLaunchKernelGGL(k_scatter_idx<chz_cl::FlagNZ>, dim3((unsigned)nb), dim3(CT_THREADS), 0, c->stream, f, ord, bc, idx);
      unsigned long long nc = 0;
      if (hipMemcpyAsync(&nc, bc + nb, 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
          hipStreamSynchronize(c->stream) != hipSuccess) { release(); chz_cl::g_err = "device error"; return 5; }
      cands[q] = (Cand *)dalloc(std::max<unsigned long long>(nc, 1) * sizeof(Cand));
      if (!cands[q]) { release(); chz_cl::g_err = "device memory"; return 5; }
      ncand[q] = (long long)nc;
      if (nc)
        hipLaunchKernelGGL(chz_cl::k_cand_build, dim3((unsigned)std::min<unsigned long long>((nc + 255) / 256, 8192)), dim3(256),
                           0, c->stream, c->sorted, idx, (long long)nc, flag, n, cands[q]);
    }
  }
  // the beam over every T gate (src/clifford.cu:318-345)
  const int grid = 4096;
  uint64_t *d_K = (uint64_t *)dalloc(chz_cl::TOPK * 8);
  int *d_cost = (int *)dalloc(chz_cl::TOPK * 4);
  unsigned long long *d_top = (unsigned long long *)dalloc((size_t)grid * chz_cl::TOPK * 8);
  if (!d_K || !d_cost || !d_top) { release(); chz_cl::g_err = "device memory"; return 5; }
  std::vector<std::vector<int>> parent(t), flips(t);
  std::vector<std::vector<Cand>> chosen(t);
  std::vector<uint64_t> hK(1, 0);
  std::vector<int> hcost(1, 0);
  for (int j = 0; j < t; j++) {
    const int q = J.tq[j];
    const long long nc = ncand[q], tot = (long long)hK.size() * nc * 2;
    if (nc == 0) { release(); chz_cl::g_err = "no placement for a T gate"; return 4; }
    CLCHK(hipMemcpyAsync(d_K, hK.data(), hK.size() * 8, hipMemcpyHostToDevice, c->stream));
    CLCHK(hipMemcpyAsync(d_cost, hcost.data(), hcost.size() * 4, hipMemcpyHostToDevice, c->stream));
    const int blocks = (int)std::min<long long>(grid, (tot + 255) / 256);
    chz_cl::BeamArgs ba{c->T, n, d_K, d_cost, (int)hK.size(), j == 0, D[j], cands[q], nc, J.td[j], beam, d_top};
    hipLaunchKernelGGL(chz_cl::k_beam, dim3(blocks), dim3(256), 0, c->stream, ba);
    std::vector<unsigned long long> top((size_t)blocks * beam);
    CLCHK(hipMemcpyAsync(top.data(), d_top, top.size() * 8, hipMemcpyDeviceToHost, c->stream));
    CLCHK(hipStreamSynchronize(c->stream));
    std::sort(top.begin(), top.end());
    const int keep = (int)std::min<long long>(beam, tot);
    parent[j].resize(keep);
    flips[j].resize(keep);
    chosen[j].resize(keep);
    std::vector<uint64_t> nk(keep);
    std::vector<int> ncost(keep);
    for (int i = 0; i < keep; i++) {
      const int b = (int)((top[i] >> 32) & 0xff);
      const long long c2 = (long long)(top[i] & 0xffffffffull);
      CLCHK(hipMemcpyAsync(&chosen[j][i], cands[q] + (c2 >> 1), sizeof(Cand), hipMemcpyDeviceToHost, c->stream));
      parent[j][i] = b;
      flips[j][i] = (int)(c2 & 1);
      ncost[i] = (int)(top[i] >> 40);
    }
    CLCHK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < keep; i++) nk[i] = chosen[j][i].K;
    hK = nk;
    hcost = ncost;
  }
  // the final segment after each kept state; the best total, first in state order
  int best = 0, bestv = 1 << 30;
  if (t > 0) {
    int *d_fin = (int *)dalloc(chz_cl::TOPK * 4);
    if (!d_fin) { release(); chz_cl::g_err = "device memory"; return 5; }
    CLCHK(hipMemcpyAsync(d_K, hK.data(), hK.size() * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(chz_cl::k_beam_final, dim3(((int)hK.size() + 63) / 64), dim3(64), 0, c->stream, c->T, n, d_K,
                       (int)hK.size(), D[t], d_fin);
    std::vector<int> fin(hK.size());
    CLCHK(hipMemcpyAsync(fin.data(), d_fin, fin.size() * 4, hipMemcpyDeviceToHost, c->stream));
    CLCHK(hipStreamSynchronize(c->stream));
    for (size_t b = 0; b < hK.size(); b++)
      if (hcost[b] + fin[b] < bestv) { bestv = hcost[b] + fin[b]; best = (int)b; }
  }
  release();
  std::vector<Cand> K(t);
  std::vector<int> flip(t), dag(t);
  for (int j = t - 1, b = best; j >= 0; j--) { K[j] = chosen[j][b]; flip[j] = flips[j][b]; b = parent[j][b]; }
  for (int j = 0; j < t; j++) {          // the T actually placed: its dagger after the sign and the flip
    const int d0 = J.td[j] ^ K[j].neg;
    if (flip[j]) K[j].Kinv = chz_cl::flipped_inv(K[j], d0, n);
    dag[j] = d0 ^ flip[j];
  }
  // the new segments K_{j-1} then D_j then K_j^-1, each read back as an optimal gate list
  std::vector<uint64_t> X(t + 1);
  for (int j = 0; j <= t; j++) {
    uint64_t x = j > 0 ? chz_cl::tcompose(K[j - 1].K, D[j], n) : D[j];
    if (j < t) x = chz_cl::tcompose(x, K[j].Kinv, n);
    X[j] = x;
  }
  constexpr int MAXW = 32;
  std::vector<uint8_t> words((size_t)(t + 1) * MAXW);
  std::vector<int> lens(t + 1);
  if (const int rc = chz_cliff_readback(c, X.data(), t + 1, MAXW, words.data(), lens.data())) return rc;
  int total = 0, orig = 0;
  for (int j = 0; j <= t; j++) {
    if (lens[j] < 0) { chz_cl::g_err = "segment not in the table"; return 4; }
    total += lens[j];
    orig += (int)J.seg[j].size();
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t0;
  char buf[160];
  std::snprintf(buf, sizeof buf, "{\"clifford_gates\": %d, \"clifford_gates_before\": %d, \"t_gates\": %d, \"seconds\": %.2f}\n",
                total, orig, t, secs);
  std::string s = buf;
  static const char *gn[] = {"h", "s", "sdg", "x", "z", "cx", "y"};
  for (int j = 0; j <= t; j++) {
    for (int i = 0; i < lens[j]; i++) {
      const int g = words[(size_t)j * MAXW + i];
      if (c->G.k[g] == 5) std::snprintf(buf, sizeof buf, "cx %d %d\n", c->G.a[g], c->G.b[g]);
      else std::snprintf(buf, sizeof buf, "%s %d\n", gn[c->G.k[g]], c->G.a[g]);
      s += buf;
    }
    if (j < t) {
      std::snprintf(buf, sizeof buf, "%s %d\n", dag[j] ? "tdg" : "t", K[j].qp);
      s += buf;
    }
  }
  if (!out || s.size() + 1 > cap) { chz_cl::g_err = "output buffer too small (" + std::to_string(s.size() + 1) + " bytes)"; return 4; }
  s.copy(out, s.size());
  out[s.size()] = 0;
  return 0;
}
