// Drivers of the gate-count search modes `--table K` (src/cheese.cu:1431-1593, :1804-1842) and `--mitm`
// (:1595-1760); kernels, storage and the method in table_join_kernels.h.  Included by search.hip inside namespace
// chz, after the batch machinery they drive.

// the gates of table entry (level l, node j) from the identity, in order
static std::vector<int> gt_sequence(const GTable &T, int l, long long j) {
  std::vector<int> w;
  for (int x = l; x >= 1; x--) {
    w.push_back(T.gate[x][j]);
    j = T.par[x][j];
  }
  std::reverse(w.begin(), w.end());
  return w;
}

// the gates of frontier node j of the context's current level (its tree walked up to the root)
static std::vector<int> walk_node(chz_ctx *c, uint32_t j) {
  std::vector<int> seq;
  uint32_t i = j;
  for (int s = c->gstep; s >= 0; s--) {
    uint32_t par;
    uint8_t g;
    copy_sync(&par, c->tree_par[s] + i, 4, hipMemcpyDeviceToHost, c->stream);
    if (par == NO_PARENT) break;
    copy_sync(&g, c->tree_gate[s] + i, 1, hipMemcpyDeviceToHost, c->stream);
    if (g != CARRY_GATE) seq.push_back(g);
    i = par;
  }
  std::reverse(seq.begin(), seq.end());
  return seq;
}

// residual of a gate sequence from a start matrix (S <- S g^-1 per gate), on the host
static HostMat host_residual(const HostSpec &hs, const std::vector<Zq> &start, int k, const std::vector<int> &seq) {
  HostMat S;
  S.e = start;
  S.k = k;
  for (int g : seq) host_apply_inverse(S, hs.gates[g], hs.n);
  return S;
}

static std::vector<Zq> identity_entries(int n) {
  const int N = 1 << n;
  std::vector<Zq> I((size_t)N * N, Zq{0, 0, 0, 0});
  for (int i = 0; i < N; i++) I[(size_t)i * N + i] = Zq{1, 0, 0, 0};
  return I;
}

// ends every search of the context's batch without a result (the table or join answered instead)
static void batch_clear(chz_ctx *c) {
  HIPCHK(hipStreamSynchronize(c->stream));
  for (BInst &b : c->binst) b.live = false;
  c->pending.clear();
  c->events.clear();
  c->forder.clear();
  c->m = 0;
  prune_tree(c);
}

// a context for an auxiliary search of the same gate set and symmetry group (table build, backward join side)
static chz_ctx *aux_context(const chz_ctx *c, const HostSpec &hs, int depth) {
  std::unique_ptr<chz_ctx> a(new chz_ctx());
  a->hs = hs;
  a->grp = c->grp;
  a->opts = c->opts;
  a->opts.table = 0;
  a->opts.mitm = 0;
  a->depth = depth;
  a->device = c->device;
  for (int b = 0; b < 32; b++) a->sched[b] = c->sched[b];
  HIPCHK(hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking));
  StreamScope scope(a->stream);
  upload_spec(a.get());
  return a.release();
}

// builds the distance table of K levels (src/cheese.cu:1483-1518): its level lines go to cb with dir 1
template <int NQ> static void gt_build(chz_ctx *c, int K, chz_level_cb cb, void *user) {
  constexpr int N = 1 << NQ, E = N * N;
  const double t0 = now_s();
  GTable *T = new GTable();
  T->K = K;
  HostSpec hs = c->hs;                          // from the identity, no bounds, never "found"
  hs.target = identity_entries(hs.n);
  hs.tk = 0;
  hs.standard = 0;
  hs.lite = 0;
  chz_ctx *t = nullptr;
  try {
    t = aux_context(c, hs, K + 1);
    t->ks.gk = -1;                              // no residual has exponent -1: no goal
    t->goal_skip = false;
    t->nogoal = true;
    StreamScope scope(t->stream);
    {   // level 0: the identity's class
      DBuf<int4> r0;
      DBuf<int> k0;
      r0.need(E);
      k0.need(1);
      std::vector<int4> I(E, make_int4(0, 0, 0, 0));
      for (int i = 0; i < N; i++) I[i * N + i] = make_int4(1, 0, 0, 0);
      const int zero = 0;
      HIPCHK(hipMemcpyAsync(r0.p, I.data(), E * sizeof(int4), hipMemcpyHostToDevice, t->stream));
      HIPCHK(hipMemcpyAsync(k0.p, &zero, 4, hipMemcpyHostToDevice, t->stream));
      gt_add_level<NQ>(t->ks, *T, r0.p, k0.p, nullptr, nullptr, 1, t->stream);
      r0.release();
      k0.release();
    }
    t->gt_filter = T;
    if (batch_add_parsed(t, hs, K + 1, 0)) throw DeviceError();
    int levels = 0;
    bool done = K == 0;
    while (!done) {
      if (!chz_batch_live(t) && t->events.empty()) break;
      batch_step<NQ>(t);
      while (!t->events.empty()) {
        chz_event e = t->events.front();
        t->events.pop_front();
        if (e.kind == 0) {
          e.level.dir = 1;
          T->lines.push_back(e.level);
          if (cb) cb(&e.level, user);
          if (++levels >= K) done = true;
        } else {
          done = true;                          // nothing survived (a level without new classes)
        }
      }
    }
    HIPCHK(hipStreamSynchronize(t->stream));
    chz_close(t);
  } catch (DeviceError &) {
    if (t) chz_close(t);
    T->release();
    delete T;
    throw;
  }
  T->seconds = now_s() - t0;
  c->gt = T;
}

// the table lookups of the current frontier (level t, budget left maxc): the first confirmed circuit in (cost,
// node) order (src/cheese.cu:1547-1590), or empty.  `root`: level 0, the target itself (no tree yet).
template <int NQ> static bool gt_lookup(chz_ctx *c, int t, int maxc, const std::vector<int> &inv,
                                        const int4 *res, const int *pk, long long m, std::vector<int> &circuit,
                                        int &cost_out) {
  constexpr int N = 1 << NQ, E = N * N, CPW = 64 / E;
  const GTable &T = *c->gt;
  if (m <= 0 || maxc < 0) return false;
  c->gt_keys.need(m);
  c->gt_cost.need(m);
  c->gt_hit.need(m);
  c->gt_best.need(1);
  const unsigned long long none = ~0ull;
  HIPCHK(hipMemcpyAsync(c->gt_best.p, &none, 8, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_node_keys<NQ>, dim3(grid_for((m + CPW - 1) / CPW)), dim3(256), 0, c->stream, c->ks, res, pk, m,
                     c->gt_keys.p);
  hipLaunchKernelGGL(k_gt_probe<NQ>, dim3(grid_for((m + CPW - 1) / CPW)), dim3(256), 0, c->stream, c->ks, T.dev(),
                     res, pk, c->gt_keys.p, m, maxc, c->gt_cost.p, c->gt_hit.p, c->gt_best.p);
  HIPCHK(hipGetLastError());
  unsigned long long best = none;
  HIPCHK(hipMemcpyAsync(&best, c->gt_best.p, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (best == none) return false;
  // candidates in (cost, node) order; almost always the first is confirmed
  std::vector<uint8_t> cost(m);
  copy_sync(cost.data(), c->gt_cost.p, (size_t)m, hipMemcpyDeviceToHost, c->stream);
  std::vector<std::pair<int, long long>> cand;
  for (long long j = 0; j < m; j++) if (cost[j] <= maxc) cand.push_back({cost[j], j});
  std::sort(cand.begin(), cand.end());
  const HostSpec &hs = c->hs;
  const std::vector<Zq> I = identity_entries(hs.n);
  for (const auto &cj : cand) {
    uint64_t at = 0;
    copy_sync(&at, c->gt_hit.p + cj.second, 8, hipMemcpyDeviceToHost, c->stream);
    const int l = (int)(at >> 32);
    const long long x = (long long)(at & 0xffffffffull);
    const std::vector<int> pre = t > 0 ? walk_node(c, (uint32_t)cj.second) : std::vector<int>();
    const std::vector<int> w = gt_sequence(T, l, x);
    const HostMat S = host_residual(hs, hs.target, hs.tk, pre);
    const HostMat Em = host_residual(hs, I, 0, w);           // the table sequence's residual W^-1
    if (S.ovf || Em.ovf || S.k != Em.k) continue;
    for (int sy = 0; sy < c->grp.size(); sy++) {              // sigma(E) == S, first element in group order
      const uint8_t *pmap = c->grp.bperm.data() + (size_t)sy * N;
      bool eq = true;
      for (int e = 0; eq && e < E; e++) {
        Zq v = Em.e[e];
        if (c->grp.conj[sy]) v = Zq{v.a, -v.d, -v.c, -v.b};
        const Zq &u = S.e[(size_t)pmap[e >> NQ] * N + pmap[e & (N - 1)]];
        eq = v.a == u.a && v.b == u.b && v.c == u.c && v.d == u.d;
      }
      if (!eq) continue;
      std::vector<int> full = pre;                             // the sequence with product E: w reversed, inverted
      for (int i = (int)w.size() - 1; i >= 0; i--) full.push_back(c->grp.gmap[sy][inv[w[i]]]);
      if (verify_hs(hs, full.data(), (int)full.size()) == 1) {
        circuit = full;
        cost_out = l;
        return true;
      }
    }
  }
  return false;
}

static void result_circuit(chz_result *res, const std::vector<int> &circ, long found, double t0) {
  res->rc = 0;
  res->len = (int)circ.size();
  res->found = found;
  for (size_t i = 0; i < circ.size() && i < CHZ_DMAX; i++) res->circuit[i] = circ[i];
  res->seconds = now_s() - t0;
}

// `--table K` (src/cheese.cu:1804-1842): table lines when the table is built, the search's levels 1..D-K with a
// lookup after each, then a circuit from the table or the proof line
template <int NQ> static int run_table(chz_ctx *c, chz_level_cb cb, void *user, chz_result *res) {
  const int K = c->opts.table, D = c->depth;
  const std::vector<int> inv = inverse_candidates(c->hs);
  for (size_t g = 0; g < inv.size(); g++)
    if (inv[g] < 0) { set_err("--table: no inverse candidate for " + c->hs.gates[g].name); return res->rc = 3; }
  {   // depth 0: the target is the goal
    HostMat U;
    U.e = c->hs.target;
    U.k = c->hs.tk;
    if (host_equal(U, c->hs.goal, c->hs.gk)) { res->rc = 0; res->len = 0; return 0; }
  }
  if (!c->gt || c->gt->K != K) {
    if (c->gt) { c->gt->release(); delete c->gt; c->gt = nullptr; }
    gt_build<NQ>(c, K, cb, user);
  }
  const double t0 = now_s();
  std::vector<int> circ;
  int cst = 0;
  {   // level 0: the target's own class
    constexpr int E = 1 << (2 * NQ);
    DBuf<int4> r0;
    DBuf<int> k0;
    r0.need(E);
    k0.need(1);
    std::vector<int4> U(E);
    for (int i = 0; i < E; i++) {
      const Zq &z = c->hs.target[i];
      U[i] = make_int4((int)z.a, (int)z.b, (int)z.c, (int)z.d);
    }
    HIPCHK(hipMemcpyAsync(r0.p, U.data(), E * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(k0.p, &c->hs.tk, 4, hipMemcpyHostToDevice, c->stream));
    const bool hit = gt_lookup<NQ>(c, 0, D, inv, r0.p, k0.p, 1, circ, cst);
    HIPCHK(hipStreamSynchronize(c->stream));
    r0.release();
    k0.release();
    if (hit) {
      result_circuit(res, circ, 1, t0);
      res->split_kind = 1; res->split[0] = 0; res->split[1] = cst;
      return 0;
    }
  }
  if (const int rc = batch_add_parsed(c, c->hs, D, 0)) return res->rc = rc;
  for (int t = 1; t <= D - K; t++) {
    batch_step<NQ>(c);
    while (!c->events.empty()) {
      const chz_event e = c->events.front();
      c->events.pop_front();
      if (e.kind == 0) { if (cb) cb(&e.level, user); }
      else { *res = e.result; res->seconds = now_s() - t0; batch_clear(c); return res->rc; }
    }
    if (gt_lookup<NQ>(c, t, D - t, inv, c->res[c->cur].p, c->pk[c->cur].p, c->m, circ, cst)) {
      batch_clear(c);
      result_circuit(res, circ, 1, t0);
      res->split_kind = 1; res->split[0] = t; res->split[1] = cst;
      return 0;
    }
  }
  batch_clear(c);
  // the last level searched was D - K: every residual outside the table needs more than K gates
  res->rc = 1;
  res->split_kind = 1; res->split[0] = -1; res->split[1] = K;
  res->seconds = now_s() - t0;
  return 1;
}

// `--mitm` (src/cheese.cu:1664-1760)
template <int NQ> static int run_mitm(chz_ctx *c, chz_level_cb cb, void *user, chz_result *res) {
  constexpr int N = 1 << NQ, E = N * N, CPW = 64 / E;
  const int D = c->depth;
  const HostSpec &hs = c->hs;
  const std::vector<int> inv = inverse_candidates(hs);
  for (size_t g = 0; g < inv.size(); g++)
    if (inv[g] < 0) { set_err("--mitm: no inverse candidate for " + hs.gates[g].name); return res->rc = 3; }
  {
    HostMat U;
    U.e = hs.target;
    U.k = hs.tk;
    if (host_equal(U, hs.goal, hs.gk)) { res->rc = 0; res->len = 0; return 0; }
  }
  const double t0 = now_s();
  HostSpec hb = hs;                                   // U^dag, the same sqrt2 exponent
  for (int r = 0; r < N; r++)
    for (int cc = 0; cc < N; cc++) {
      const Zq &z = hs.target[(size_t)cc * N + r];
      hb.target[(size_t)r * N + cc] = Zq{z.a, -z.d, -z.c, -z.b};
    }
  chz_ctx *bk = aux_context(c, hb, D);
  struct Level { DBuf<int4> res; DBuf<int> pk; long long n = 0; };
  std::vector<std::unique_ptr<Level>> blev;           // backward levels 1..b, resident for the join
  GTable F;                                           // the forward level's multimap
  F.owns_levels = false;
  auto finish = [&](int rc) {
    batch_clear(c);
    {
      StreamScope sb(bk->stream);
      batch_clear(bk);
      for (auto &l : blev) { l->res.release(); l->pk.release(); }
      chz_close(bk);
    }
    F.release();
    return res->rc = rc;
  };
  try {
    if (const int rc = batch_add_parsed(c, hs, D, 0)) return finish(rc);
    {
      StreamScope sb(bk->stream);
      if (const int rc = batch_add_parsed(bk, hb, D, 0)) return finish(rc);
    }
    int a = 0, b = 0;
    long long nfa = 1, nbb = 1;
    bool bdead = false;
    while (a + b < D) {
      const bool fwd = bdead || nfa <= nbb;
      chz_ctx *x = fwd ? c : bk;
      StreamScope sx(x->stream);
      batch_step<NQ>(x);
      bool dead = false;
      while (!x->events.empty()) {
        chz_event e = x->events.front();
        x->events.pop_front();
        if (e.kind == 0) {
          e.level.dir = fwd ? 2 : 3;
          if (cb) cb(&e.level, user);
          continue;
        }
        if (e.result.rc == 0) {
          std::vector<int> circ(e.result.circuit, e.result.circuit + e.result.len);
          if (!fwd) {                                 // a circuit q for U^dag: U = q^-1
            std::vector<int> rev;
            for (int i = (int)circ.size() - 1; i >= 0; i--) rev.push_back(inv[circ[i]]);
            circ = rev;
          }
          result_circuit(res, circ, e.result.found, t0);
          finish(0);
          return 0;
        }
        if (e.result.rc != 1 || fwd) {                // overflow, or a dead forward side: the search's answer
          const int rc = e.result.rc;
          finish(rc);
          res->seconds = now_s() - t0;
          return rc;
        }
        dead = true;                                  // a dead backward side: no suffix longer than b
      }
      if (dead) { bdead = true; break; }
      if (fwd) { a++; nfa = c->m; }
      else {
        b++;
        nbb = bk->m;
        std::unique_ptr<Level> l(new Level());
        l->n = bk->m;
        l->res.need((size_t)std::max<long long>(l->n, 1) * E);
        l->pk.need(std::max<long long>(l->n, 1));
        if (l->n) {
          HIPCHK(hipMemcpyAsync(l->res.p, bk->res[bk->cur].p, (size_t)l->n * E * sizeof(int4), hipMemcpyDeviceToDevice, bk->stream));
          HIPCHK(hipMemcpyAsync(l->pk.p, bk->pk[bk->cur].p, (size_t)l->n * 4, hipMemcpyDeviceToDevice, bk->stream));
        }
        HIPCHK(hipStreamSynchronize(bk->stream));
        blev.push_back(std::move(l));
      }
    }
    const double tj = now_s();
    long long nb = 0;
    for (auto &l : blev) nb += l->n;
    if (!nb) {
      res->rc = 1;
      res->split_kind = 2; res->split[0] = a; res->split[1] = b;
      finish(1);
      res->seconds = now_s() - t0;
      return 1;
    }
    StreamScope sc(c->stream);
    const long long nf = c->m;                        // forward level a (a >= 1: the forward side moves first)
    gt_add_level<NQ>(c->ks, F, c->res[c->cur].p, c->pk[c->cur].p, nullptr, nullptr, nf, c->stream);
    DBuf<int4> dU;
    dU.need(E);
    std::vector<int4> Uh(E);
    for (int i = 0; i < E; i++) Uh[i] = make_int4((int)hs.target[i].a, (int)hs.target[i].b, (int)hs.target[i].c, (int)hs.target[i].d);
    HIPCHK(hipMemcpyAsync(dU.p, Uh.data(), E * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    DBuf<uint8_t> hashes;
    DBuf<long long> fidx;
    DBuf<unsigned long long> ovf;
    hashes.need(nb);
    fidx.need(nb);
    ovf.need(1);
    HIPCHK(hipMemsetAsync(ovf.p, 0, 8, c->stream));
    long long off = 0;
    for (auto &l : blev) {
      if (l->n)
        hipLaunchKernelGGL(k_mitm_join<NQ>, dim3(grid_for((l->n + CPW - 1) / CPW)), dim3(256), 0, c->stream, c->ks, F.dev(),
                           dU.p, hs.tk, l->res.p, l->pk.p, l->n, hashes.p + off, fidx.p + off, ovf.p);
      off += l->n;
    }
    HIPCHK(hipGetLastError());
    std::vector<uint8_t> hh(nb);
    std::vector<long long> ff(nb);
    unsigned long long novf = 0;
    HIPCHK(hipMemcpyAsync(hh.data(), hashes.p, (size_t)nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(ff.data(), fidx.p, (size_t)nb * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&novf, ovf.p, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dU.release(); hashes.release(); fidx.release(); ovf.release();
    if (novf) { set_err("--mitm: coefficients exceed int32"); finish(4); return 4; }
    long long nh = 0;
    for (long long j = 0; j < nb; j++) nh += hh[j] != 0;
    {   // the join line
      chz_level jl{};
      jl.dir = 4;
      jl.parents = (long)nf;
      jl.prefixes = (long)nb;
      jl.distinct = (long)nh;
      jl.seconds = now_s() - tj;
      if (cb) cb(&jl, user);
    }
    off = 0;
    for (size_t li = 0; li < blev.size(); li++) {
      for (long long j = 0; j < blev[li]->n; j++) {
        if (ff[off + j] < 0) continue;
        // backward node j of level li + 1: its sequence q (the suffix reversed and inverted)
        std::vector<int> q;
        {
          uint32_t i = (uint32_t)j;
          for (int s = (int)li + 1; s >= 1; s--) {
            uint32_t par;
            uint8_t g;
            copy_sync(&par, bk->tree_par[s] + i, 4, hipMemcpyDeviceToHost, bk->stream);
            copy_sync(&g, bk->tree_gate[s] + i, 1, hipMemcpyDeviceToHost, bk->stream);
            q.push_back(g);
            i = par;
          }
          std::reverse(q.begin(), q.end());
        }
        std::vector<int> circ = walk_node(c, (uint32_t)ff[off + j]);
        for (int i = (int)q.size() - 1; i >= 0; i--) circ.push_back(inv[q[i]]);
        if (verify_hs(hs, circ.data(), (int)circ.size()) != 1) continue;
        result_circuit(res, circ, 1, t0);
        res->split_kind = 2; res->split[0] = a; res->split[1] = (int)q.size();
        finish(0);
        return 0;
      }
      off += blev[li]->n;
    }
    res->rc = 1;
    res->split_kind = 2; res->split[0] = a; res->split[1] = b;
    finish(1);
    res->seconds = now_s() - t0;
    return 1;
  } catch (DeviceError &) {
    try { finish(5); } catch (DeviceError &) {}
    return res->rc = 5;
  }
}
