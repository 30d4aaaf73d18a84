// Request server over the C ABI (no reference counterpart: the reference answers one request per process).
// Requests of three kinds arrive from any thread: gate-count synthesis (.spec text and budget), T-count
// (.tspec text and maximum), Clifford re-synthesis (a gate-set word whose optimal word is wanted).
//
// Served mode is continuous batching.  One worker thread per gate-count batch (2 and 3 qubits) steps its batch a
// level at a time and admits queued searches between steps while the estimated frontier memory of the batch
// stays within the budget; a T-count worker searches every queued target together; a Clifford worker answers
// every queued word in one batched read-back.  Each worker owns its contexts and streams, so their kernels
// overlap on the device.  Plain mode serves one request at a time, in arrival order, on the same engines (a
// gate-count search as a batch of one), which is the single-instance path the served mode is measured against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cheese_hip.h"
#include "host.h"

namespace {

struct Req {
  chz_request r;
  std::string text;
  std::vector<uint8_t> word;
  double t_submit;
};

double clock_s(const std::chrono::steady_clock::time_point &t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct chz_server {
  chz_server_opts o{};
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req> q_gates[2], q_tc, q_cl, q_plain;   // gate queues by qubit count (2, 3)
  std::deque<chz_response> done;
  std::atomic<long long> outstanding{0};
  std::atomic<bool> stop{false};
  std::vector<std::thread> workers;
  // engines
  std::string base[2];          // gate-count base specs (2 and 3 qubits)
  chz_ctx *gctx[2] = {nullptr, nullptr};
  chz_tc *tc = nullptr;
  chz_cliff *cl = nullptr;
  std::string err;
  int plain = 0;
  bool dead[2] = {false, false};   // a gate engine that could not reopen after a device error (answers rc 5)
  std::atomic<long long> mem_peak{0};       // largest frontier-memory estimate of a gate batch
  std::atomic<long long> used_peak{0};      // largest device memory in use, sampled after every gate step
  // per-engine statistics (under mu): gate batches [2], T-count, Clifford
  double steps[2] = {0, 0}, step_s[2] = {0, 0}, kernel_ms[2] = {0, 0}, live_peak[2] = {0, 0};
  double tc_batches = 0, tc_s = 0, cl_batches = 0, cl_s = 0;
};

namespace {

void finish(chz_server *s, const chz_response &resp) {
  std::lock_guard<std::mutex> lk(s->mu);
  s->done.push_back(resp);
  s->outstanding--;
}

void raise_to(std::atomic<long long> &a, long long v) {
  long long cur = a.load();
  while (v > cur && !a.compare_exchange_weak(cur, v)) {}
}

// device memory in use now (every process's), 0 if unknown; taken under the capture lock (a device-wide query)
long long device_used_bytes() {
  size_t free_b = 0, total_b = 0;
  std::lock_guard<std::recursive_mutex> g(chz::capture_mu());
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
  return (long long)(total_b - free_b);
}

// kernel families an engine times for its statistics: the expansion and the captured level steps, which hold
// nearly all of a batch's kernel time, at one event pair per launch
constexpr unsigned kServerTimed = (1u << CHZ_FAM_EXPAND) | (1u << CHZ_FAM_LEVEL_GRAPH);

double ctx_kernel_ms(const chz_ctx *c) {
  chz_kstat ks[16];
  const int n = chz_kernel_stats(c, ks, 16);
  double t = 0;
  for (int i = 0; i < n; i++) t += ks[i].ms;
  return t;
}

chz_response base_resp(chz_server *s, const Req &q) {
  chz_response r{};
  r.id = q.r.id;
  r.kind = q.r.kind;
  r.t_submit = q.t_submit;
  (void)s;
  return r;
}

// Memory: each gate-count batch gets a share of the budget (kShare), enforced level by level by the batch's own
// planning (chz_batch_set_budget: searches that would not fit wait a step, carried unexpanded).  Admission keeps
// the frontier a new search joins within half of that share: per frontier node its residual and the children a
// level makes (kNodeBytes, kChildBytes), a root charged for its first level of G children, at most kAdmitPerStep
// joining one step.  kMaxLive is a throughput setting, not a memory bound: the 2,000-request burst is served
// fastest with about 64 live 3-qubit searches per batch.
constexpr double kNodeBytes[2] = {16.0 * 16 + 24, 16.0 * 64 + 24};
constexpr double kChildBytes = 48.0;
constexpr double kShare[2] = {0.2, 0.8};
constexpr int kAdmitPerStep = 64;
constexpr int kMaxLive[2] = {1024, 64};

double per_node_bytes(int qi, int G) { return kNodeBytes[qi] + G * kChildBytes; }

// ---- gate-count searches: one batch per qubit count, stepped level by level
void gate_worker(chz_server *s, int qi) {
  chz::RelaxedCapture relaxed;    // this thread's calls never invalidate another engine's capture
  chz_ctx *ctx = s->gctx[qi];
  const int G = chz_num_gates(ctx);
  const double share = s->o.mem_budget_gb * 1e9 * kShare[qi];
  chz_batch_set_budget(ctx, share);
  std::map<int, Req> live;      // tag -> request
  int next_tag = 0;
  std::vector<chz_event> ev(256);
  while (true) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (live.empty())
        s->cv.wait_for(lk, std::chrono::milliseconds(50), [&] { return s->stop || !s->q_gates[qi].empty(); });
      if (s->stop && live.empty() && s->q_gates[qi].empty()) return;
      // admission: queued searches join while the frontier the batch holds, with every search admitted in this
      // round counted at once, stays within half the batch's share
      const double per_node = per_node_bytes(qi, G);
      double held = (double)chz_batch_nodes(ctx) * per_node;
      const double want = (double)G * per_node;
      int admitted = 0;
      while (!s->q_gates[qi].empty()) {
        if (!live.empty() && (held + want > 0.5 * share || admitted >= kAdmitPerStep)) break;
        if ((int)live.size() >= (s->o.max_batch > 0 ? s->o.max_batch : kMaxLive[qi])) break;
        held += want;
        admitted++;
        Req q = std::move(s->q_gates[qi].front());
        s->q_gates[qi].pop_front();
        lk.unlock();
        const int tag = next_tag++;
        const int rc = chz_batch_add(ctx, q.text.data(), q.text.size(), q.r.budget, tag);
        if (rc) {
          chz_response r = base_resp(s, q);
          r.rc = rc;
          r.t_done = clock_s(s->t0);
          finish(s, r);
        } else {
          live.emplace(tag, std::move(q));
        }
        lk.lock();
      }
    }
    if (chz_batch_live(ctx)) {
      const double ts = clock_s(s->t0);
      const int nlive = chz_batch_live(ctx);
      if (chz_batch_step(ctx)) {   // a device error: every live search of this batch ends with it
        {
          std::lock_guard<std::mutex> lk(s->mu);
          s->err = chz_last_error();
        }
        for (auto &kv : live) {
          chz_response r = base_resp(s, kv.second);
          r.rc = 5;
          r.t_done = clock_s(s->t0);
          finish(s, r);
        }
        live.clear();
        // queued searches go on in a fresh context; if none can be opened they end with the error too
        chz_close(ctx);
        ctx = s->gctx[qi] = nullptr;
        chz_opts co{-1, 1, 1, 0, 0, s->o.device};
        chz_ctx *fresh = nullptr;
        if (chz_open(s->base[qi].data(), s->base[qi].size(), &co, &fresh) == 0) {
          ctx = s->gctx[qi] = fresh;
          chz_time_kernels(ctx, kServerTimed);
          chz_batch_set_budget(ctx, share);
          continue;
        }
        // no engine left for this qubit count: queued and later requests are answered with the error at once
        std::lock_guard<std::mutex> lk(s->mu);
        s->dead[qi] = true;
        while (!s->q_gates[qi].empty()) {
          chz_response r = base_resp(s, s->q_gates[qi].front());
          r.rc = 5;
          r.t_done = clock_s(s->t0);
          s->done.push_back(r);
          s->outstanding--;
          s->q_gates[qi].pop_front();
        }
        return;
      }
      raise_to(s->mem_peak, (long long)((double)chz_batch_nodes(ctx) * per_node_bytes(qi, G)));
      raise_to(s->used_peak, device_used_bytes());
      const double kms = ctx_kernel_ms(ctx);
      std::lock_guard<std::mutex> lk(s->mu);
      s->steps[qi] += 1;
      s->step_s[qi] += clock_s(s->t0) - ts;
      s->kernel_ms[qi] = kms;
      s->live_peak[qi] = std::max(s->live_peak[qi], (double)nlive);
    }
    int n;
    while ((n = chz_batch_events(ctx, ev.data(), (int)ev.size())) > 0)
      for (int i = 0; i < n; i++) {
        if (ev[i].kind != 1) continue;
        auto it = live.find(ev[i].tag);
        if (it == live.end()) continue;
        chz_response r = base_resp(s, it->second);
        r.rc = ev[i].result.rc;
        r.len = ev[i].result.rc == 0 ? ev[i].result.len : 0;
        for (int j = 0; j < r.len && j < CHZ_DMAX; j++) r.out[j] = ev[i].result.circuit[j];
        r.found = ev[i].result.found;
        r.t_done = clock_s(s->t0);
        live.erase(it);
        finish(s, r);
      }
  }
}

void tc_fill(const chz_tc_result &t, chz_response &r) {
  r.rc = t.rc;
  r.len = t.rc == 0 ? t.tcount : 0;
  for (int k = 0; k < r.len; k++) { r.out[2 * k] = t.rotations[2 * k]; r.out[2 * k + 1] = t.rotations[2 * k + 1]; }
  r.split[0] = t.split[0];
  r.split[1] = t.split[1];
}

// ---- T-count: every queued target in one batched search
void tc_worker(chz_server *s) {
  chz::RelaxedCapture relaxed;    // this thread's calls never invalidate another engine's capture
  while (true) {
    std::vector<Req> batch;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait_for(lk, std::chrono::milliseconds(50), [&] { return s->stop || !s->q_tc.empty(); });
      if (s->stop && s->q_tc.empty()) return;
      const int cap = s->o.max_batch > 0 ? s->o.max_batch : 1024;
      while (!s->q_tc.empty() && (int)batch.size() < cap) { batch.push_back(std::move(s->q_tc.front())); s->q_tc.pop_front(); }
    }
    if (batch.empty()) continue;
    const double ts = clock_s(s->t0);
    // one search per maximum (the maximum bounds how deep a batch peels)
    std::map<int, std::vector<int>> bymax;
    for (int i = 0; i < (int)batch.size(); i++) bymax[batch[i].r.budget >= 0 ? batch[i].r.budget : 16].push_back(i);
    for (auto &kv : bymax) {
      const int nb = (int)kv.second.size();
      std::vector<const char *> sp(nb);
      std::vector<size_t> ln(nb);
      std::vector<chz_tc_result> res(nb);
      for (int j = 0; j < nb; j++) { sp[j] = batch[kv.second[j]].text.data(); ln[j] = batch[kv.second[j]].text.size(); }
      const int rc = chz_tc_search(s->tc, sp.data(), ln.data(), nb, kv.first, res.data());
      const double t = clock_s(s->t0);
      for (int j = 0; j < nb; j++) {
        chz_response r = base_resp(s, batch[kv.second[j]]);
        if (rc) r.rc = rc;
        else tc_fill(res[j], r);
        r.t_done = t;
        finish(s, r);
      }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->tc_batches += 1;
    s->tc_s += clock_s(s->t0) - ts;
  }
}

void cl_fill(chz_server *s, std::vector<Req> &batch) {
  const long long n = (long long)batch.size();
  constexpr int MAXW = 32;
  std::vector<uint64_t> tab(n);
  for (long long i = 0; i < n; i++)
    tab[i] = chz_cliff_word_tableau(s->cl, batch[i].word.data(), (int)batch[i].word.size());
  std::vector<uint8_t> words(n * MAXW);
  std::vector<int> lens(n);
  const int rc = chz_cliff_readback(s->cl, tab.data(), n, MAXW, words.data(), lens.data());
  const double t = clock_s(s->t0);
  for (long long i = 0; i < n; i++) {
    chz_response r = base_resp(s, batch[i]);
    r.rc = rc ? rc : (lens[i] < 0 ? 4 : 0);
    r.len = rc || lens[i] < 0 ? 0 : lens[i];
    for (int j = 0; j < r.len; j++) r.out[j] = words[i * MAXW + j];
    r.t_done = t;
    finish(s, r);
  }
}

// ---- Clifford re-synthesis: every queued word in one read-back launch
void cl_worker(chz_server *s) {
  chz::RelaxedCapture relaxed;    // this thread's calls never invalidate another engine's capture
  while (true) {
    std::vector<Req> batch;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait_for(lk, std::chrono::milliseconds(50), [&] { return s->stop || !s->q_cl.empty(); });
      if (s->stop && s->q_cl.empty()) return;
      while (!s->q_cl.empty()) { batch.push_back(std::move(s->q_cl.front())); s->q_cl.pop_front(); }
    }
    if (batch.empty()) continue;
    const double ts = clock_s(s->t0);
    cl_fill(s, batch);
    std::lock_guard<std::mutex> lk(s->mu);
    s->cl_batches += 1;
    s->cl_s += clock_s(s->t0) - ts;
  }
}

// ---- plain: one request at a time, in arrival order
void plain_worker(chz_server *s) {
  chz::RelaxedCapture relaxed;    // this thread's calls never invalidate another engine's capture
  std::vector<chz_event> ev(64);
  while (true) {
    Req q;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait_for(lk, std::chrono::milliseconds(50), [&] { return s->stop || !s->q_plain.empty(); });
      if (s->q_plain.empty()) {
        if (s->stop) return;
        continue;
      }
      q = std::move(s->q_plain.front());
      s->q_plain.pop_front();
    }
    chz_response r = base_resp(s, q);
    if (q.r.kind == CHZ_REQ_GATES) {
      chz_ctx *ctx = nullptr;
      for (int qi = 0; qi < 2; qi++) if (s->gctx[qi] && chz_batch_add(s->gctx[qi], q.text.data(), q.text.size(), q.r.budget, 0) == 0) { ctx = s->gctx[qi]; break; }
      if (!ctx) r.rc = 3;
      bool got = false;
      while (ctx && !got) {
        int n;
        while ((n = chz_batch_events(ctx, ev.data(), (int)ev.size())) > 0)
          for (int i = 0; i < n; i++)
            if (ev[i].kind == 1) {
              r.rc = ev[i].result.rc;
              r.len = ev[i].result.rc == 0 ? ev[i].result.len : 0;
              for (int j = 0; j < r.len; j++) r.out[j] = ev[i].result.circuit[j];
              r.found = ev[i].result.found;
              got = true;
            }
        if (!got && chz_batch_step(ctx)) { r.rc = 5; got = true; }
      }
    } else if (q.r.kind == CHZ_REQ_TCOUNT && !s->tc) {
      r.rc = 3;
    } else if (q.r.kind == CHZ_REQ_TCOUNT) {
      const char *sp = q.text.data();
      const size_t ln = q.text.size();
      chz_tc_result res{};
      const int rc = chz_tc_search(s->tc, &sp, &ln, 1, q.r.budget >= 0 ? q.r.budget : 16, &res);
      if (rc) r.rc = rc;
      else tc_fill(res, r);
    } else {
      std::vector<Req> one;
      one.push_back(std::move(q));
      cl_fill(s, one);
      continue;
    }
    r.t_done = clock_s(s->t0);
    finish(s, r);
  }
}

}  // namespace

extern "C" {

int chz_server_open(const char *spec2, size_t len2, const char *spec3, size_t len3, const char *tspec3, size_t lent,
                    const chz_server_opts *opts, chz_server **out) {
  *out = nullptr;
  chz_server *s = new chz_server();
  if (opts) s->o = *opts;
  if (s->o.mem_budget_gb <= 0) s->o.mem_budget_gb = 200.0;
  if (s->o.tcount_K <= 0) s->o.tcount_K = 4;
  s->plain = s->o.plain;
  chz_opts co{-1, 1, 1, 0, 0, s->o.device};
  const char *bs[2] = {spec2, spec3};
  const size_t bl[2] = {len2, len3};
  for (int qi = 0; qi < 2; qi++) {
    if (!bs[qi]) continue;
    s->base[qi].assign(bs[qi], bl[qi]);
    if (const int rc = chz_open(bs[qi], bl[qi], &co, &s->gctx[qi])) { chz_server_close(s); return rc; }
    chz_time_kernels(s->gctx[qi], kServerTimed);
  }
  if (tspec3) {
    if (const int rc = chz_tc_table(tspec3, lent, s->o.tcount_K, s->o.device, &s->tc)) { chz_server_close(s); return rc; }
  }
  if (const int rc = chz_cliff_build(3, s->o.device, &s->cl)) { chz_server_close(s); return rc; }
  s->t0 = std::chrono::steady_clock::now();
  if (s->plain) {
    s->workers.emplace_back(plain_worker, s);
  } else {
    for (int qi = 0; qi < 2; qi++) if (s->gctx[qi]) s->workers.emplace_back(gate_worker, s, qi);
    if (s->tc) s->workers.emplace_back(tc_worker, s);
    s->workers.emplace_back(cl_worker, s);
  }
  *out = s;
  return 0;
}

double chz_server_clock(const chz_server *s) { return clock_s(s->t0); }

int chz_server_submit(chz_server *s, const chz_request *r) {
  Req q;
  q.r = *r;
  if (r->text && r->len) q.text.assign(r->text, r->len);
  if (r->word && r->wlen > 0) q.word.assign(r->word, r->word + r->wlen);
  q.r.text = nullptr;
  q.r.word = nullptr;
  q.t_submit = clock_s(s->t0);
  if (r->kind != CHZ_REQ_GATES && r->kind != CHZ_REQ_TCOUNT && r->kind != CHZ_REQ_CLIFFORD) return 3;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    // a request no engine can take is answered at once: rc 3 when its kind or qubit count is not served, rc 5
    // when its engine ended after a device error
    int now_rc = 0;
    int qi = -1;
    if (r->kind == CHZ_REQ_GATES) {
      // route by qubit count: the "n" line of the spec
      const size_t p = q.text.find("\nn ");
      const int n = p == std::string::npos ? 0 : std::atoi(q.text.c_str() + p + 3);
      qi = n == 2 ? 0 : n == 3 ? 1 : -1;
      if (qi < 0 || !s->gctx[qi]) now_rc = 3;
      else if (s->dead[qi]) now_rc = 5;
    } else if (r->kind == CHZ_REQ_TCOUNT && !s->tc) {
      now_rc = 3;
    }
    s->outstanding++;
    if (now_rc) {
      chz_response a = base_resp(s, q);
      a.rc = now_rc;
      a.t_done = q.t_submit;
      s->done.push_back(a);
      s->outstanding--;
      return 0;
    }
    if (s->plain) s->q_plain.push_back(std::move(q));
    else if (r->kind == CHZ_REQ_GATES) s->q_gates[qi].push_back(std::move(q));
    else if (r->kind == CHZ_REQ_TCOUNT) s->q_tc.push_back(std::move(q));
    else s->q_cl.push_back(std::move(q));
  }
  s->cv.notify_all();
  return 0;
}

int chz_server_poll(chz_server *s, chz_response *out, int max) {
  std::lock_guard<std::mutex> lk(s->mu);
  int k = 0;
  while (k < max && !s->done.empty()) { out[k++] = s->done.front(); s->done.pop_front(); }
  return k;
}

long long chz_server_outstanding(const chz_server *s) { return s->outstanding.load(); }

long long chz_server_peak_bytes(const chz_server *s) { return s->mem_peak.load(); }

long long chz_server_peak_device_bytes(const chz_server *s) { return s->used_peak.load(); }

int chz_server_stats(chz_server *s, double *out, int n) {
  std::lock_guard<std::mutex> lk(s->mu);
  const double v[] = {s->steps[0], s->steps[1], s->step_s[0], s->step_s[1], s->kernel_ms[0], s->kernel_ms[1],
                      s->live_peak[0], s->live_peak[1], s->tc_batches, s->tc_s, s->cl_batches, s->cl_s};
  const int k = std::min(n, (int)(sizeof v / sizeof v[0]));
  for (int i = 0; i < k; i++) out[i] = v[i];
  return k;
}

const char *chz_server_error(chz_server *s) {
  std::lock_guard<std::mutex> lk(s->mu);
  return s->err.c_str();
}

void chz_server_close(chz_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto &t : s->workers) t.join();
  for (int qi = 0; qi < 2; qi++) if (s->gctx[qi]) chz_close(s->gctx[qi]);
  if (s->tc) chz_tc_free(s->tc);
  if (s->cl) chz_cliff_free(s->cl);
  delete s;
}

}  // extern "C"
