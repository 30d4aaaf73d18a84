// Gate-count search on one MI355X: per-level expansion, T bound, ordered compaction, LDS-staged LSD radix
// sort, segmented exact dedupe and frontier gather, all hand-written HIP kernels for gfx950, driven by the
// C-ABI in include/cheese_hip.h.  Reference behaviour: src/cheese.cu:1249-1377 (levels), :1762-1848 (run).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cheese_hip.h"
#include "host.h"
#include "primitives.h"
#include "search_kernels.h"
#include "spec.h"

static int verify_hs(const chz::HostSpec &hs, const int *circuit, int len);

namespace chz {

// ------------------------------------------------------------------------------------------------------
// counters of one level (device): indices into a u64 array
// C_EVAL children whose residual is formed, C_CHEAP through the structural bounds, C_CUTS cut ranks computed,
// C_NEEDT T-bound evaluations (their f = 0 parts), C_KEYS class keys, C_TPASS flip-pair passes of the T bound: the
// stage counts of the expansion's work model
// C_M, C_NCC: the level's parents and children, written by the host with the level's control upload (read by the
// kernels of a captured level step, whose launches are sized for a capacity)
enum { C_W2 = 0, C_W3, C_DEAD, C_FOUND, C_ALIVE, C_OVF, C_NEEDT, C_FIRST, C_RUNS, C_REPOVF, C_KEPT, C_EVAL, C_SURV,
       C_CHEAP, C_CUTS, C_KEYS, C_TPASS, C_M, C_NCC, C_N };
// per-instance counters of one level (a batch carries several searches that share the spec's gate set)
enum { IC_W2 = 0, IC_W3, IC_DEAD, IC_FOUND, IC_ALIVE, IC_OVF, IC_FIRST, IC_N };

struct ExpandArgs {
  const int4 *res;          // parents' residuals, E entries each
  const int *pk;            // parents' sqrt2 exponents
  const uint8_t *g1, *g2;   // parents' last / second-to-last gates (255: none)
  long long off;            // first parent of the chunk (index in the level)
  long long ncc;            // children in the chunk
  int t, keep;
  uint8_t *status;          // [ncc]
  uint64_t *keys;           // [ncc] class keys of kept children
  uint32_t *info;           // [ncc] bound words of kept children
  const uint32_t *pb;       // parents' bound words
  unsigned long long *cnt;  // global counters (stage counts C_EVAL, C_CHEAP, C_CUTS, C_NEEDT, C_KEYS, C_TPASS)
  const uint32_t *pinst;    // instance slot of every parent of the level (null: one instance, slot 0)
  const int *rem;           // per slot: gate budget left after this level's gate (depth - t)
  const uint8_t *keepi;     // per slot: this level's survivors become parents (t < depth)
  unsigned long long *icnt; // per slot: IC_N counters
  int gsplit;               // a parent's candidates are split among this many lane groups (small levels)
  const unsigned long long *npd;   // parents read on the device (captured step; gsplit then from tw), or null
  int tw;                   // candidate-split threshold in waves (with npd)
  const uint32_t *gflag;    // per slot: some parent of the level has a goal child (null: not tracked)
};

// One level's children, fused (src/cheese.cu:662-720, :959-1021): a lane group per parent holds the parent's
// residual in LDS and walks all G candidates: windows, child residual, overflow, goal, cheap bounds, T bound
// (mod 2^16, one lane per flip pair, when n + k <= 16; else mod 2^32 / 2^64, one lane per entry) and the
// class key of a kept child.  Writes the status byte (and key) of every child of the chunk.
// kind: 0 general, 1 unit entries, 2 unit permutation; tdel, loc, cdel: DevGate::tdel, ::cutloc, ::cdel
struct GateHdr { int nq, L, mask, ik, q0, q1, q2, kind, tdel, loc, cdel, cls; };
template <int NQ> __device__ __forceinline__ GateHdr wuni(const GateHdr &h) {
  return GateHdr{wuni<NQ>(h.nq), wuni<NQ>(h.L), wuni<NQ>(h.mask), wuni<NQ>(h.ik), wuni<NQ>(h.q0), wuni<NQ>(h.q1),
                 wuni<NQ>(h.q2), wuni<NQ>(h.kind), wuni<NQ>(h.tdel), wuni<NQ>(h.loc), wuni<NQ>(h.cdel),
                 wuni<NQ>(h.cls)};
}
constexpr int MAXG = 64;

// The expansion kernel's LDS tables, built once per context on the host (build_level_tabs) and copied into LDS by
// every block with coalesced 16-byte loads; built inside the kernel they cost each block scattered global loads,
// which small levels (one parent per wave) cannot amortise.
struct LevelTabs {
  // per candidate and entry (g * 64 + e), the child entry's terms for unit gates with at most two per column:
  // bits 0-2 source column x0, 3-5 power p0 of w, 6-8 x1, 9-11 p1, bit 12 two terms; bit 15: use the table loop
  uint16_t sdesc[MAXG * 64];
  GateHdr shdr[MAXG];
  uint8_t spos[16 * 64];           // n = 3: position of entry e under symmetry y (y * 64 + e)
  uint8_t scx[CXL_MAX];            // the cx table when it fits (KSpec::cxlds)
  uint8_t scsrc[CSRC_MAX * 64];    // cut i's operator-Schmidt matrix: entry source of lane le (i * GS + le)
};
static_assert(sizeof(LevelTabs) % 16 == 0, "LevelTabs is copied as uint4");

template <int NQ, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_level(KSpec sp, ExpandArgs a,
                                                                                         long long np) {
  constexpr int N = 1 << NQ, E = N * N, GS = E, CPW = 64 / GS;
  __shared__ uint2 stage_x[TS_STRIDE];
  __shared__ uint4 stage_p[2 * TS_STRIDE];
  __shared__ int4 spar[256];                   // each group's parent residual
  __shared__ __align__(16) LevelTabs tb;
  for (int i = threadIdx.x; i < (int)(sizeof(LevelTabs) / 16); i += blockDim.x) ((uint4 *)&tb)[i] = sp.ktab[i];
  const uint16_t *const sdesc = tb.sdesc;
  const GateHdr *const shdr = tb.shdr;
  const uint8_t *const spos = tb.spos, *const scx = tb.scx, *const scsrc = tb.scsrc;
  const uint32_t cjm = NQ == 3 ? sp.cjm : 0u;
  __syncthreads();
  const int lane = threadIdx.x & 63, le = lane & (GS - 1), slot = lane / GS;
  const int r = le >> NQ, c = le & (N - 1);
  const TStage st{stage_x + (threadIdx.x & ~(GS - 1)), stage_p + (threadIdx.x & ~(GS - 1))};
  int4 *P = spar + (threadIdx.x & ~(GS - 1));
  const long long wave = wave_index();
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  const int G = sp.G;
  // per-instance status tallies: one LDS counter per status value and group, bumped by the group's leader lane
  // and flushed when the group's instance changes (32-bit: one wave's share of a launch is far below 2^32
  // children).  Held in registers they cost two vector or scalar instructions per tally and child, and the
  // scalar registers they take push others into spill lanes.
  __shared__ uint32_t stally[256 / GS][8];
  uint32_t *const tl = stally[threadIdx.x / GS];
  for (int i = le; i < 8; i += GS) tl[i] = 0;
  __builtin_amdgcn_wave_barrier();
  uint32_t n_eval = 0, n_t = 0, n_cheap = 0, n_cuts = 0, n_keys = 0, n_tp = 0;
  int cur = -1;
  auto flush = [&]() {
    __builtin_amdgcn_wave_barrier();
    if (le == 0 && cur >= 0) {
      unsigned long long *ic = a.icnt + (size_t)cur * IC_N;
      const uint32_t w2 = tl[ST_W2], w3 = tl[ST_W3], dead = tl[ST_DEAD], found = tl[ST_FOUND], ovf = tl[ST_OVF],
                     alive = tl[ST_ALIVE];
      if (w2) atomicAdd(&ic[IC_W2], (unsigned long long)w2);
      if (w3) atomicAdd(&ic[IC_W3], (unsigned long long)w3);
      if (dead) atomicAdd(&ic[IC_DEAD], (unsigned long long)dead);
      if (found) atomicAdd(&ic[IC_FOUND], (unsigned long long)found);
      if (ovf) atomicAdd(&ic[IC_OVF], (unsigned long long)ovf);
      if (alive) atomicAdd(&ic[IC_ALIVE], (unsigned long long)alive);
    }
    __builtin_amdgcn_wave_barrier();
    for (int i = le; i < 8; i += GS) tl[i] = 0;
    __builtin_amdgcn_wave_barrier();
  };
  int gsplit = a.gsplit;
  if (a.npd) {
    np = (long long)*a.npd;
    const long long gs = np ? ((long long)a.tw * CPW + np - 1) / np : 1;
    gsplit = (int)(gs < 1 ? 1 : gs > G ? G : gs);
  }
  const long long nitems = np * gsplit;
  const int cs = (G + gsplit - 1) / gsplit;
  for (long long base = wave * CPW; base < nitems; base += nwaves * CPW) {
    const long long item = base + slot;
    if (item >= nitems) continue;
    // the 64-bit division runs on the vector unit; readfirstlane keeps the parent and the candidate range (the
    // child loop's bounds) scalar for n = 3
    const long long pl = wuni64<NQ>(item / gsplit);
    const int gbeg = wuni<NQ>((int)(item - pl * gsplit) * cs), gend = gbeg + cs < G ? gbeg + cs : G;
    const long long pi = a.off + pl;
    const int inst = wuni<NQ>(a.pinst ? (int)a.pinst[pi] : 0);
    if (inst != cur) { flush(); cur = inst; }
    const int rem = wuni<NQ>(a.rem[inst]);
    const bool keep = wuni<NQ>((int)a.keepi[inst]) != 0;
    const int4 S = a.res[pi * E + le];
    const int pk = wuni<NQ>(a.pk[pi]), g1 = wuni<NQ>((int)a.g1[pi]), g2 = wuni<NQ>((int)a.g2[pi]);
    const uint32_t pbw = (uint32_t)wuni<NQ>((int)a.pb[pi]);
    const bool small = gall<GS>(zmag(S) < (1u << 28));
    // A search whose circuit is found at this level ends with it: its parents without a goal child (no
    // BW_GOAL) are passed over once their coefficients rule out an overflowing child.  Goal children come only
    // from BW_GOAL parents, so the level's found count, first circuit and overflow test are unchanged.
    if (a.gflag && !(pbw & BW_GOAL) && wuni<NQ>((int)a.gflag[inst]) != 0 && gall<GS>(zmag(S) <= sp.maglim)) {
      for (int g = gbeg + le; g < gend; g += GS) a.status[pl * G + g] = ST_DEAD;
      continue;
    }
    __builtin_amdgcn_wave_barrier();
    P[le] = S;
    __builtin_amdgcn_wave_barrier();
    // window cuts of every candidate of this parent, one candidate per lane, as two bit masks: the candidate
    // loop then tests bits instead of waiting on a table load per child
    // lower cut bound of each cut-delta class (one per parent instead of one per child)
    const bool cdcache = !sp.lite && sp.cx_mono && (pbw & BW_CUTS) && sp.ncd > 0;
    uint64_t c2c = 0;
    if (cdcache)
      for (int j = 0; j < sp.ncd; j++) {
        const int v = cx_lower(sp, scx, (pbw & BW_NEEDS) >> 8, sp.cdv[j]);
        c2c |= (uint64_t)(v < 255 ? v : 255) << (8 * j);
      }
    // T bounds of this parent's children per T-bound class: tc_ok bit set when one is known, tc_ex when it is the
    // exact bound, else it is a lower bound (a transform stopped once the child was dead)
    uint64_t tc_val = 0;
    uint32_t tc_ok = 0, tc_ex = 0;
    uint64_t w2m = 0, w3m = 0;
    for (int b0 = 0; b0 < G; b0 += GS) {
      const int g = b0 + le;
      bool c2 = false, c3 = false;
      if (g < G && g1 != 255) {
        c2 = sp.bad2[g1 * G + g] != 0;
        if (g2 != 255) {
          const uint32_t tri = ((uint32_t)g2 * G + g1) * G + g;
          c3 = ((sp.bad3[tri >> 5] >> (tri & 31)) & 1) != 0;
        }
      }
      w2m |= gballot<GS>(c2) << b0;
      w3m |= gballot<GS>(c3) << b0;
    }
    for (int g = gbeg; g < gend; g++) {
      const long long cc = pl * G + g;
      uint8_t stt = 255;
      if ((w2m >> g) & 1) stt = ST_W2;
      else if ((w3m >> g) & 1) stt = ST_W3;
      uint64_t key = 0;
      uint32_t bw = BW_TUNK;
      if (stt == 255) {
        // child residual entry (r, c) = sum_m P[r][x_m] ginv[m][lc] (src/cheese.cu:146-174)
        const GateHdr h = wuni<NQ>(shdr[g]);
        int lc = (c >> h.q0) & 1;
        if (h.nq > 1) lc |= ((c >> h.q1) & 1) << 1;
        if (h.nq > 2) lc |= ((c >> h.q2) & 1) << 2;
        const int cb = c & ~h.mask;
        const int8_t *ph = sp.gates[g].ph;    // table loop only: gates with more than two terms per column
        int4 s;
        bool of = false;
        const uint32_t desc = sdesc[g * 64 + le];
        if (h.kind && small && !(desc & 0x8000u)) {
          s = zrot(P[r * N + (desc & 7)], (desc >> 3) & 7);
          if (desc & 0x1000u) {
            const int4 t = zrot(P[r * N + ((desc >> 6) & 7)], (desc >> 9) & 7);
            s.x += t.x; s.y += t.y; s.z += t.z; s.w += t.w;
          }
        } else if (h.kind && small) {
          int4 acc = make_int4(0, 0, 0, 0);
          for (int m = 0; m < h.L; m++) {
            const int p = ph[m * h.L + lc];
            if (p < 0) continue;
            int x = cb | ((m & 1) << h.q0);
            if (h.nq > 1) x |= ((m >> 1) & 1) << h.q1;
            if (h.nq > 2) x |= ((m >> 2) & 1) << h.q2;
            const int4 t = zrot(P[r * N + x], p);
            acc.x += t.x; acc.y += t.y; acc.z += t.z; acc.w += t.w;
            if (h.kind == 2) break;
          }
          s = acc;
        } else {
          const DevGate &dg = sp.gates[g];
          long long va = 0, vb = 0, vc = 0, vd = 0;
          for (int m = 0; m < h.L; m++) {
            const int4 gm = dg.inv[m * h.L + lc];
            if (zzero(gm)) continue;
            int x = cb | ((m & 1) << h.q0);
            if (h.nq > 1) x |= ((m >> 1) & 1) << h.q1;
            if (h.nq > 2) x |= ((m >> 2) & 1) << h.q2;
            const int4 v = P[r * N + x];
            va += (long long)v.x * gm.x - (long long)v.y * gm.w - (long long)v.z * gm.z - (long long)v.w * gm.y;
            vb += (long long)v.x * gm.y + (long long)v.y * gm.x - (long long)v.z * gm.w - (long long)v.w * gm.z;
            vc += (long long)v.x * gm.z + (long long)v.y * gm.y + (long long)v.z * gm.x - (long long)v.w * gm.w;
            vd += (long long)v.x * gm.w + (long long)v.y * gm.z + (long long)v.z * gm.y + (long long)v.w * gm.x;
          }
          of = va != (int)va || vb != (int)vb || vc != (int)vc || vd != (int)vd;
          s = make_int4((int)va, (int)vb, (int)vc, (int)vd);
        }
        int k = pk + h.ik;
        n_eval++;
        if (gany<GS>(of)) stt = ST_OVF;
        else {
          while (k > 0 && gall<GS>(zhalf(s))) { s = zdown(s); k--; }
          // the goal's entries are read only for a child with the goal's exponent and nonzero pattern
          if (k == sp.gk && gballot<GS>(!zzero(s)) == sp.goalnz && gall<GS>(zeq(s, sp.goal[le]))) stt = ST_FOUND;
          else {
            bool alive = true;
            if (sp.standard || sp.lite) {
              int hh, supp = 0, c2 = 0;
              if (sp.lite) hh = k;
              else {
                n_cheap++;
                const uint64_t nzb = gballot<GS>(!zzero(s));
                const int hs = h_struct<NQ>(nzb);
                hh = k > hs ? k : hs;
                supp = support<NQ>(s, le, nzb);
              }
              // Bracket the child's bounds from its parent's: the T bound moves by at most the gate's transfer-
              // matrix exponent h.tdel (Clifford gates keep it), each cut need drops by at most h.cdel.  The
              // bound combination is monotone in T and in the cut bound, so the brackets settle most children
              // before the cut ranks and the T transform are computed.
              const bool tpar = (pbw & BW_TUNK) != BW_TUNK && h.tdel != 255;
              const int Tp = (int)(pbw & BW_TUNK);
              // a sibling of the same T-bound class already has its exact bound (it is this child's too) or a
              // lower bound
              const int tc = (h.cls >> 8) - 1;
              const bool tcached = tc >= 0 && ((tc_ok >> tc) & 1);
              const int tcv = tcached ? (int)((tc_val >> (8 * tc)) & 0xff) : 0;
              const bool tcex = tcached && ((tc_ex >> tc) & 1);
              const bool tknown = (tpar && h.tdel == 0) || tcex;
              const int Tbr = tpar ? (Tp > h.tdel ? Tp - h.tdel : 0) : 0;
              int T = tcex ? tcv : (Tbr > tcv ? Tbr : tcv);    // exact if tknown, else a lower bound
              bool texact = tknown;
              const int c2lo = cdcache ? (int)((c2c >> (8 * (h.cls & 0xff))) & 0xff)
                               : !sp.lite && sp.cx_mono && (pbw & BW_CUTS) ? cx_lower(sp, scx, (pbw & BW_NEEDS) >> 8, (uint32_t)h.cdel)
                               : 0;
              if (combine_bound(hh, (T + sp.tdiv - 1) / sp.tdiv, c2lo, supp) > rem) alive = false;
              else {
                if (!sp.lite) {
                  uint32_t needs = (pbw & BW_NEEDS) >> 8;
                  const uint32_t keepc = (pbw & BW_CUTS) ? (uint32_t)h.loc : 0u;
                  n_cuts += sp.ncuts - __popc(keepc & ((1u << sp.ncuts) - 1));
                  c2 = wuni<NQ>(cx_bound<NQ>(s, le, sp, scx, scsrc, keepc, needs));
                  bw = (needs << 8 & ~BW_CUTS) | (sp.ncuts <= BW_MAXCUTS ? BW_CUTS : 0u) | BW_TUNK;
                }
                if (combine_bound(hh, (T + sp.tdiv - 1) / sp.tdiv, c2, supp) > rem) alive = false;
                else if (!texact && tpar && rem <= 1 &&
                         combine_bound(hh, (Tp + h.tdel + sp.tdiv - 1) / sp.tdiv, c2, supp) <= rem) {
                  // alive whatever the exact T; its children (budget 0) need no inherited T bound
                  alive = true;
                } else if (!texact) {
                  n_t++;
                  if (NQ + k <= 16) {
                    // the f = 0 part, then the flip-pair passes (odd first) while the lower bound leaves the child
                    // alive: at T bound T it is alive exactly when T <= tmax
                    uint32_t np = 0;
                    T = t_bound16s<NQ>(st, s, k, le, T, t_budget(hh, c2, supp, rem) * sp.tdiv, np, texact);
                    n_tp += np;
                  } else {
                    T = t_bound_any<NQ>(s, k, le);
                    texact = true;
                  }
                  if (tc >= 0 && T < 255 && (texact || !tcached || T > tcv)) {
                    tc_val = (tc_val & ~(0xffull << (8 * tc))) | (uint64_t)T << (8 * tc);
                    tc_ok |= 1u << tc;
                    if (texact) tc_ex |= 1u << tc;
                  }
                  alive = combine_bound(hh, (T + sp.tdiv - 1) / sp.tdiv, c2, supp) <= rem;
                }
              }
              if (alive && texact && T < (int)BW_TUNK) bw = (bw & ~BW_TUNK) | (uint32_t)T;
            }
            if (alive && keep) {
              n_keys++;
              if constexpr (NQ == 3) {
                key = sp.nsym <= 16 ? class_key3(s, k, lane, spos, cjm, sp.nsym) : class_key<NQ>(s, k, le, sp);
              } else {
                key = class_key<NQ>(s, k, le, sp);
              }
              key ^= (uint64_t)inst * 0x9e3779b97f4a7c15ull;
            }
            stt = alive ? ST_ALIVE : ST_DEAD;
          }
        }
      }
      if (le == 0) {
        atomicAdd(&tl[stt], 1u);
        a.status[cc] = stt;
        if (stt == ST_ALIVE && keep) { a.keys[cc] = key; a.info[cc] = bw; }
        if (stt == ST_FOUND)
          atomicMin(&a.icnt[(size_t)inst * IC_N + IC_FIRST], (unsigned long long)(a.off * G + cc));
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  flush();
  if (le != 0) n_eval = n_t = n_cheap = n_cuts = n_keys = n_tp = 0;     // one copy per group
  n_eval = gsum<64>(n_eval);
  n_t = gsum<64>(n_t);
  n_cheap = gsum<64>(n_cheap);
  n_cuts = gsum<64>(n_cuts);
  n_keys = gsum<64>(n_keys);
  n_tp = gsum<64>(n_tp);
  if (lane == 0) {
    if (n_eval) atomicAdd(&a.cnt[C_EVAL], (unsigned long long)n_eval);
    if (n_t) atomicAdd(&a.cnt[C_NEEDT], (unsigned long long)n_t);
    if (n_cheap) atomicAdd(&a.cnt[C_CHEAP], (unsigned long long)n_cheap);
    if (n_cuts) atomicAdd(&a.cnt[C_CUTS], (unsigned long long)n_cuts);
    if (n_keys) atomicAdd(&a.cnt[C_KEYS], (unsigned long long)n_keys);
    if (n_tp) atomicAdd(&a.cnt[C_TPASS], (unsigned long long)n_tp);
  }
}

struct FlagAlive {      // survivors of a chunk that become parents, in child order
  const uint8_t *status;
  const uint32_t *pinst;
  const uint8_t *keepi;
  long long off;          // first parent of the chunk
  int G;
  __device__ bool operator()(long long i) const {
    return status[i] == ST_ALIVE && keepi[pinst ? pinst[off + i / G] : 0] != 0;
  }
};
struct FlagKept {       // kept survivors of instances still searching
  const uint8_t *kept;
  const uint64_t *surv_o;
  const uint32_t *pinst;
  const uint8_t *active;
  int G;
  __device__ bool operator()(long long i) const {
    return kept[i] != 0 && active[pinst ? pinst[surv_o[i] / (uint64_t)G] : 0] != 0;
  }
};

// survivors of a chunk -> level survivor list (child index in the level, class key)
__global__ __launch_bounds__(CT_THREADS) void k_scatter_alive(FlagAlive f, long long n, const unsigned long long *boff,
                                                              const uint64_t *keys, long long child0,
                                                              const unsigned long long *out0, uint64_t *surv_o,
                                                              uint64_t *surv_key, const uint32_t *info,
                                                              uint32_t *surv_bw, const unsigned long long *nd) {
  n = dev_n(n, nd);
  __shared__ int ws[CT_THREADS / 64];
  const long long t0 = (long long)blockIdx.x * CT_TILE;
  long long pos = (long long)*out0 + (long long)boff[blockIdx.x];
  for (int i = 0; i < CT_ITEMS; i++) {
    const long long x = t0 + (long long)i * CT_THREADS + threadIdx.x;
    const bool fl = x < n && f(x);
    int ex;
    const int tot = block_flag_scan(fl, ex, ws);
    if (fl) {
      surv_o[pos + ex] = (uint64_t)(child0 + x);
      surv_key[pos + ex] = keys[x];
      surv_bw[pos + ex] = info[x];
    }
    pos += tot;
  }
}

// the same for a chunk of one tile (n <= CT_TILE) in one launch: the level's survivor count *out0 is read at the
// start and advanced at the end
__global__ __launch_bounds__(CT_THREADS) void k_scatter_alive_one(FlagAlive f, long long n, const uint64_t *keys,
                                                                  long long child0, unsigned long long *out0,
                                                                  uint64_t *surv_o, uint64_t *surv_key,
                                                                  const uint32_t *info, uint32_t *surv_bw) {
  __shared__ int ws[CT_THREADS / 64];
  const long long p0 = (long long)*out0;
  long long pos = p0;
  for (int i = 0; i < CT_ITEMS; i++) {
    const long long x = (long long)i * CT_THREADS + threadIdx.x;
    const bool fl = x < n && f(x);
    int ex;
    const int tot = block_flag_scan(fl, ex, ws);
    if (fl) {
      surv_o[pos + ex] = (uint64_t)(child0 + x);
      surv_key[pos + ex] = keys[x];
      surv_bw[pos + ex] = info[x];
    }
    pos += tot;
  }
  if (threadIdx.x == 0) *out0 = (unsigned long long)pos;
}

// searches that continue after this level: no circuit, no overflow, budget left (captured step)
__global__ void k_active(const unsigned long long *icnt, const uint8_t *keepi, int nslots, uint8_t *active) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots; s += gridDim.x * blockDim.x) {
    const unsigned long long *x = icnt + (size_t)s * IC_N;
    active[s] = !x[IC_FOUND] && !x[IC_OVF] && keepi[s];
  }
}

// the level's survivor count grows by a chunk's total (the scan's last entry)
__global__ void k_add_total(unsigned long long *acc, const unsigned long long *tot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *acc += *tot;
}

// ------------------------------------------------------------------------------------------------------
// segmented dedupe over the sorted survivors: runs of equal 32-bit key prefix are partitioned into exact
// classes, keeping the smallest survivor rank (= lexicographically first prefix) of each class
// (src/cheese.cu:1334-1343, :1046-1065)
// every survivor starts kept (a sequential fill: setting kept[v[i]] here would scatter byte writes over the array)
// (it runs after the level's expansion has read the goal flags, before the gather sets them for the next level:
// gflag, when given, is cleared here)
__global__ void k_fill_kept(uint8_t *kept, long long n, const unsigned long long *nd, uint32_t *gflag, int ns) {
  n = dev_n(n, nd);
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x, stride = (long long)gridDim.x * blockDim.x;
  for (long long i = i0; i < n; i += stride) kept[i] = 1;
  if (gflag)
    for (long long i = i0; i < ns; i += stride) gflag[i] = 0u;
}

struct DedupeArgs {
  const uint32_t *k, *v;          // sorted key prefixes and survivor ranks
  long long n;
  const uint64_t *surv_o, *surv_key;
  const uint32_t *runs;
  const int4 *res;                // parents' residuals
  const int *pk;
  uint8_t *kept;
  unsigned long long *cnt;
  const uint32_t *pinst;          // parents' instance slots (null: one instance)
  const unsigned long long *nd;   // survivors read on the device (captured step), or null
};
constexpr int MAXREP = 32, MAXSYM = 32;

// value of lane q of the lane group (q uniform across the group)
template <int GS> __device__ __forceinline__ uint32_t gget(uint32_t v, int q) {
  if constexpr (GS == 64) return (uint32_t)__builtin_amdgcn_readlane((int)v, q);
  else return (uint32_t)__shfl((int)v, q, GS);
}
template <int GS> __device__ __forceinline__ uint64_t gget64(uint64_t v, int q) {
  return (uint64_t)gget<GS>((uint32_t)v, q) | ((uint64_t)gget<GS>((uint32_t)(v >> 32), q) << 32);
}

template <int NQ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NQ > 1 ? 6 : 1)))
void k_dedupe_runs(KSpec sp, DedupeArgs a) {
  constexpr int N = 1 << NQ, E = N * N, GS = E, CPW = 64 / GS;
  constexpr int GPB = 256 / GS;     // groups per block
  __shared__ uint64_t rep_key[GPB][MAXREP], rep_o[GPB][MAXREP];
  __shared__ uint32_t rep_inst[GPB][MAXREP];
  __shared__ uint8_t spos[MAXSYM * E];
  __shared__ int4 rep_res[GPB][2][E];     // the first two representatives' residuals
  __shared__ int4 par_stage[GPB][E];      // the current member's parent residual
  // the symmetry group's entry maps, staged once per block (the class test tries the elements in order)
  const int nsym = sp.nsym < MAXSYM ? sp.nsym : MAXSYM;
  const bool lds_sym = nsym == sp.nsym;
  uint32_t cjm = 0;
  for (int i = threadIdx.x; i < nsym * E; i += blockDim.x) {
    const int y = i / E, e = i % E;
    spos[i] = (uint8_t)(sp.symperm[y * N + (e >> NQ)] * N + sp.symperm[y * N + (e & (N - 1))]);
  }
  for (int y = 0; y < nsym; y++) cjm |= (uint32_t)(sp.symconj[y] != 0) << y;
  __syncthreads();
  const int lane = threadIdx.x & 63, le = lane & (GS - 1), slot = lane / GS;
  const int grp = (threadIdx.x >> 6) * CPW + slot;
  const long long wave = wave_index();
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  const long long nruns = (long long)a.cnt[C_RUNS];
  const long long n = dev_n(a.n, a.nd);
  const uint64_t G = (uint64_t)sp.G;
  for (long long base = wave * CPW; base < nruns; base += nwaves * CPW) {
    const long long ri = base + slot;
    if (ri >= nruns) continue;
    const long long i0 = a.runs[ri];
    const uint32_t key = a.k[i0];
    int nrep = 0;
    // a run is almost always one class: its members are each rebuilt once and compared with the first
    // representative, whose residual (and the second's) stays in LDS
    int ck0 = 0, ck1 = 0;
    // members in chunks of GS: lane le loads member j0 + le's rank, child, key, parent exponent and slot at once,
    // so each member then waits on one load (its parent's residual), issued while the member before it is
    // compared
    for (long long j0 = i0; j0 < n; j0 += GS) {
      const long long jj = j0 + le;
      const bool in = jj < n && a.k[jj] == key;
      const uint64_t inb = gballot<GS>(in);
      const int nm = ~inb ? __ffsll((long long)~inb) - 1 : GS;     // members are contiguous from j0
      uint32_t m_rank = 0, m_inst = 0;
      uint64_t m_o = 0, m_key = 0;
      int m_pk = 0;
      if (le < nm) {
        m_rank = a.v[jj];
        m_o = a.surv_o[m_rank];
        m_key = a.surv_key[m_rank];
        const long long pi = (long long)(m_o / G);
        m_pk = a.pk[pi];
        m_inst = a.pinst ? a.pinst[pi] : 0u;
      }
      int4 nextv = make_int4(0, 0, 0, 0);
      if (nm > 0) nextv = a.res[(long long)(gget64<GS>(m_o, 0) / G) * E + le];
      for (int q = 0; q < nm; q++) {
        const uint32_t rank = gget<GS>(m_rank, q), inst = gget<GS>(m_inst, q);
        const uint64_t o = gget64<GS>(m_o, q), key64 = gget64<GS>(m_key, q);
        const int pk = (int)gget<GS>((uint32_t)m_pk, q);
        const int g = (int)(o % G);
        __builtin_amdgcn_wave_barrier();
        par_stage[grp][le] = nextv;
        __builtin_amdgcn_wave_barrier();
        if (q + 1 < nm) nextv = a.res[(long long)(gget64<GS>(m_o, q + 1) / G) * E + le];
        int4 s;
        int ks;
        make_child<NQ>(par_stage[grp], pk, sp.gates[g], le, s, ks);
        bool dup = false;
        for (int r = 0; r < nrep && !dup; r++) {
          if (rep_key[grp][r] != key64 || rep_inst[grp][r] != inst) continue;
          if (r < 2 && lds_sym) {
            dup = same_class_rep<NQ>(s, ks, rep_res[grp][r], r ? ck1 : ck0, le, spos, cjm, nsym);
          } else {
            int4 b;
            int kb;
            if (r < 2) { b = rep_res[grp][r][le]; kb = r ? ck1 : ck0; }
            else {
              const uint64_t ro = rep_o[grp][r];
              const long long rpi = (long long)(ro / G);
              make_child<NQ>(a.res + rpi * E, a.pk[rpi], sp.gates[(int)(ro % G)], le, b, kb);
            }
            dup = lds_sym ? same_class_lds<NQ>(s, ks, b, kb, le, spos, cjm, nsym) : same_class<NQ>(s, ks, b, kb, le, sp);
          }
        }
        if (dup) { if (le == 0) a.kept[rank] = 0; }
        else if (nrep < MAXREP) {
          if (le == 0) {
            rep_key[grp][nrep] = key64;
            rep_o[grp][nrep] = o;
            rep_inst[grp][nrep] = inst;
          }
          if (nrep < 2) {
            rep_res[grp][nrep][le] = s;
            if (nrep == 0) ck0 = ks;
            else ck1 = ks;
          }
          nrep++;
        } else if (le == 0) atomicAdd(&a.cnt[C_REPOVF], 1ull);
        __builtin_amdgcn_wave_barrier();
      }
      if (nm < GS) break;      // the run ended inside this chunk
    }
  }
}

struct GatherArgs {
  const uint32_t *kidx;           // kept survivor ranks, in order
  long long nk;
  const uint64_t *surv_o;
  const int4 *res;
  const int *pk;
  const uint8_t *g1;
  int4 *res_next;
  int *pk_next;
  uint8_t *g1_next, *g2_next;
  uint32_t *tree_par;
  uint8_t *tree_gate;
  const uint32_t *pinst;          // parents' instance slots (null: one instance)
  uint32_t *inst_next;
  const uint32_t *surv_bw;        // bound words of the survivors
  uint32_t *pb_next;
  const unsigned long long *nkd;  // kept count read on the device (captured step), or null
  uint32_t *gflag;                // per slot: set when a kept child has a goal child (null: not tracked)
  int tn;                         // kept nodes per wave tile (a multiple of the nodes per wave, at most 64)
};

// next frontier: tree node, window state and materialised residual of every kept child.  A wave takes tiles
// of a.tn consecutive kept nodes: one lane per node resolves the node's index chain (kept rank -> survivor ->
// parent and gate) and writes its small fields, then the lane groups build the tile's residuals from the
// parents, with the chain already in registers.
template <int NQ>
__global__ __launch_bounds__(256) void k_gather(KSpec sp, GatherArgs a) {
  constexpr int N = 1 << NQ, E = N * N, GS = E, CPW = 64 / GS;
  const int lane = threadIdx.x & 63, le = lane & (GS - 1), slot = lane / GS;
  const long long wave = wave_index();
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  const long long nk = dev_n(a.nk, a.nkd);
  const int tn = a.tn;
  for (long long tb = wave * tn; tb < nk; tb += nwaves * tn) {
    const long long pl = tb + lane;
    uint32_t pi_l = 0, bw_l = 0;
    int g_l = 0, gp_l = 255;
    if (lane < tn && pl < nk) {
      const uint32_t kid = a.kidx[pl];
      const uint64_t o = a.surv_o[kid];
      pi_l = (uint32_t)(o / (uint64_t)sp.G);
      g_l = (int)(o % (uint64_t)sp.G);
      bw_l = a.surv_bw[kid];
      gp_l = a.g1[pi_l];
      a.g1_next[pl] = (uint8_t)g_l;
      a.g2_next[pl] = (uint8_t)gp_l;
      a.tree_par[pl] = pi_l;
      a.tree_gate[pl] = (uint8_t)g_l;
      a.inst_next[pl] = a.pinst ? a.pinst[pi_l] : 0u;
    }
    for (int j = 0; j < tn; j += CPW) {
      const int src = j + slot;
      const long long pi = (long long)(uint32_t)__shfl((int)pi_l, src);
      const int g = __shfl(g_l, src), gp = __shfl(gp_l, src);
      const uint32_t bwp = (uint32_t)__shfl((int)bw_l, src);
      const long long p = tb + src;
      if (src >= tn || p >= nk) continue;
      int4 s;
      int k;
      make_child<NQ>(a.res + pi * E, a.pk[pi], sp.gates[g], le, s, k);
      a.res_next[p * E + le] = s;
      // goal child: the node equals the pre-image of a candidate its windows allow.  One lane per candidate
      // filters by exponent and nonzero pattern; the few that pass are compared exactly.
      bool adj = false;
      if (a.gflag) {
        const uint64_t nzb = gballot<GS>(!zzero(s));
        for (int b0 = 0; b0 < sp.G && !adj; b0 += GS) {
          const int x = b0 + le;
          uint64_t cm = gballot<GS>(x < sp.G && sp.xgk[x] == k && sp.xgm[x] == nzb);
          while (cm && !adj) {
            const int xx = b0 + __builtin_ctzll(cm);
            cm &= cm - 1;
            bool ok = sp.bad2[g * sp.G + xx] == 0;
            if (ok && gp != 255) {
              const uint32_t tri = ((uint32_t)gp * sp.G + g) * sp.G + xx;
              ok = ((sp.bad3[tri >> 5] >> (tri & 31)) & 1) == 0;
            }
            if (ok) adj = gall<GS>(zeq(s, sp.xg[xx * E + le]));
          }
        }
      }
      if (le == 0) {
        a.pk_next[p] = k;
        a.pb_next[p] = bwp | (adj ? BW_GOAL : 0u);
        if (adj) a.gflag[a.pinst ? a.pinst[pi] : 0u] = 1u;
      }
    }
  }
}

// nodes per gather tile: one per lane group while the launch has enough waves for all of them, up to 64
template <int NQ> static int gather_tile(long long nk) {
  constexpr int CPW = 64 / (1 << (2 * NQ));
  const long long waves = 8192ll * 4;
  long long tn = (nk + waves - 1) / waves;
  tn = (tn + CPW - 1) / CPW * CPW;
  return (int)std::max<long long>(CPW, std::min<long long>(64, tn));
}


// 64-bit class keys -> sortable 32-bit prefixes and survivor ranks
__global__ void k_init_sort(const uint64_t *key, long long n, uint32_t *k32, uint32_t *v32,
                            const unsigned long long *nd) {
  n = dev_n(n, nd);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    k32[i] = (uint32_t)(key[i] >> 32);
    v32[i] = (uint32_t)i;
  }
}

// ------------------------------------------------------------------------------------------------------
// host side

// K_STEP: a captured level step (expansion through kept compaction in one graph launch, small levels)
enum { K_EXPAND = 0, K_COMPACT, K_SORT, K_DEDUPE, K_GATHER, K_STEP, K_NFAM };
static const char *k_names[K_NFAM] = {"expand", "compact", "sort", "dedupe", "gather", "level_graph"};

#include "table_join_kernels.h"

}  // namespace chz

namespace chz {
// one search of a batch (searches that share a spec's gate set advance level by level together)
struct BInst {
  int tag = 0, depth = 0, t = 0, slot = 0, root_step = -1;
  bool live = false;
  long long nodes = 0;
  double growth = 0;           // the last level's distinct / parents (0: not known yet)
  double t0 = 0;
};
struct Pending {               // a search waiting to join the frontier as a root
  int inst, k;
  std::vector<int4> U;
};
}  // namespace chz

struct chz_ctx {
  chz::HostSpec hs;
  chz::SymGroup grp;
  chz_opts opts{};
  int depth = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  chz::KSpec ks{};
  // spec on the device
  chz::DBuf<uint8_t> spec_blob;     // the device tables below, one block
  chz::DBuf<uint8_t> ktab;          // the expansion kernel's LDS tables (LevelTabs)
  chz::DevGate *d_gates = nullptr;
  uint8_t *d_bad2 = nullptr, *d_symperm = nullptr, *d_symconj = nullptr;
  uint32_t *d_bad3 = nullptr;
  int4 *d_goal = nullptr, *d_xg = nullptr;
  uint64_t *d_xgm = nullptr;
  int *d_cx = nullptr, *d_xgk = nullptr;
  // frontier (current and next generation) and the tree of all levels
  chz::DBuf<int4> res[2];
  chz::DBuf<int> pk[2];
  chz::DBuf<uint8_t> g1[2], g2[2];
  chz::DBuf<uint32_t> pb[2];        // bound words (T bound, cut needs) of the frontier's nodes
  std::vector<uint32_t *> tree_par;
  std::vector<uint8_t *> tree_gate;
  // per-level work arrays
  chz::DBuf<uint8_t> status, kept;
  chz::DBuf<uint64_t> keys, surv_o, surv_key;
  chz::DBuf<uint32_t> info, surv_bw;
  chz::DBuf<unsigned long long> bcount, hist;
  chz::DBuf<uint32_t> sk[2], sv[2], runs, kidx;
  // batch state: searches, frontier instance slots, per-slot device arrays, queued events
  std::vector<chz::BInst> binst;
  std::vector<chz::Pending> pending;
  std::deque<chz_event> events;
  chz::DBuf<uint32_t> finst[2];
  long long m = 0;             // nodes in the current frontier (buffer cur)
  int cur = 0, gstep = -1, nslots = 0;
  bool shared_steps = false;      // the batch has held more than one search at once (captured steps are used)
  std::vector<uint32_t> cdv;      // distinct cut-delta values of the spec's gates (bits 0-7 of DevGate::cls index them)
  // level control block, uploaded in one copy at the start of every level and read back in one copy:
  // [cnt: C_N counters | icnt: per-slot IC_N counters | d_rem: per-slot budgets | d_keepi: per-slot keep flags]
  chz::DBuf<unsigned long long> ctl;
  unsigned long long *cnt = nullptr, *icnt = nullptr;
  int *d_rem = nullptr;
  uint8_t *d_keepi = nullptr;
  chz::DBuf<uint8_t> d_active;
  chz::DBuf<uint32_t> gflag;        // per slot: a node of the current frontier has a goal child
  chz::DBuf<unsigned long long> scan_tmp;
  chz::Pinned pin;                  // staging of the per-level control transfers
  // captured level steps (one per frontier parity) and the buffer addresses and schedule they were captured with
  struct StepGraph { hipGraphExec_t exec = nullptr; std::vector<uintptr_t> key; };
  StepGraph sgraph[2];
  chz::DBuf<long long> d_lo, d_hi;
  // expansion schedule per level-size bucket (log2 of the chunk's parents): occupancy variant, grid cap in
  // blocks, candidate-split threshold in waves; chosen by chz_tune, defaults otherwise
  struct Sched { int var = 1, cap = 8192, tw = 4096; };
  Sched sched[32];
  bool tuning = false;
  Sched force;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> tune_ev;   // (bucket, events) per launch
  std::map<void *, size_t> tree_blocks;   // every tree block allocated, with its size
  std::vector<void *> tree_free;          // blocks no live search reaches, ready for reuse
  // measurement
  chz::EventLog ev;
  long k_launch[chz::K_NFAM] = {};
  double k_ms[chz::K_NFAM] = {}, k_work[chz::K_NFAM] = {};
  double k_stage[6] = {};       // expansion stage counts: residuals, structural bounds, cut ranks, T bounds, keys,
                                // T-bound passes
  double avg_terms = 1;         // nonzero terms per inverse column, averaged over the candidates
  bool goal_skip = false;       // every candidate has a goal pre-image: the found-level skip is exact
  // per-level memory planning (chz_batch_set_budget): searches in frontier order; with a budget, a level expands
  // the longest prefix of them whose planned bytes fit, and the rest wait, carried to the next frontier unexpanded
  std::vector<int> forder;
  double mem_budget = 0;
  long long carried_levels = 0;   // search-levels that waited (measurement)
  // --table K: the distance table (built on first use, kept with the context); a context that builds one has
  // gt_filter set (its levels drop classes the table holds and are appended to it) and no goal test
  chz::GTable *gt = nullptr;
  chz::GTable *gt_filter = nullptr;
  bool nogoal = false;
  chz::DBuf<uint64_t> gt_keys, gt_hit;
  chz::DBuf<uint8_t> gt_cost;
  chz::DBuf<unsigned long long> gt_best;
};

namespace chz {

// (the context's stream has synchronised: the blocks go back to the process-wide cache)
static void free_tree(chz_ctx *c) {
  for (auto &kv : c->tree_blocks) dev_pool().put(kv.first, kv.second);
  c->tree_blocks.clear();
  c->tree_free.clear();
  c->tree_par.clear();
  c->tree_gate.clear();
}

// one tree level of n nodes (parents u32, then gates u8) from the block pool
static void tree_alloc(chz_ctx *c, long long n, uint32_t *&tp, uint8_t *&tg) {
  const size_t need = (size_t)std::max<long long>(n, 1) * 5;
  void *blk = nullptr;
  for (size_t i = 0; i < c->tree_free.size(); i++)
    if (c->tree_blocks[c->tree_free[i]] >= need) {
      blk = c->tree_free[i];
      c->tree_free.erase(c->tree_free.begin() + i);
      break;
    }
  if (!blk) {
    size_t got = 0;
    blk = dev_pool().get(need + need / 4 + 256, got);
    c->tree_blocks[blk] = got;
  }
  tp = (uint32_t *)blk;
  tg = (uint8_t *)(tp + std::max<long long>(n, 1));
}

// Lane operations of the T bound as implemented.  The f = 0 part (32-bit): per entry 12 for |S|^2 = r + y sqrt2
// and 2 per transform stage for each of r and y.  One flip-pair pass (16-bit, two to a packed lane operation): for
// each of the 4^n - 1 flip pairs, half of E/2 Z[w] products (16 multiplies + 12 adds), of a (2n-1)-stage
// transform of E/2 four-coefficient entries and of 6 operations per entry for the final combination.
static double t_f0_ops(int nq) {
  const double E = (double)(1 << (2 * nq));
  return E * 12 + 2 * nq * E * 2;
}
static double t_pass_ops(int nq) {
  const double E = (double)(1 << (2 * nq)), H = E / 2;
  return (E - 1) * (H * 28 + (2 * nq - 1) * H * 4 + H * 6) / 2 / 2;
}

// Algorithmic integer operations of one level's expansion, by stage, in 32-bit lane operations (two 16-bit
// operations of the packed T transform count as one):
//   residual: per entry, 8 per nonzero term of the gate's inverse column (4 coefficient rotations, 4 adds), plus
//             6 for the sqrt2-divisibility test and 4 for the goal comparison;
//   structural bounds: per entry 4 for the nonzero test, 5 per qubit for the support comparison;
//   cut rank: per entry 30 for the F_p image (4 Montgomery products and reductions) and 14 per elimination
//             step (2 Montgomery products, 1 modular subtraction), stop = 3 steps;
//   T bound: t_f0_ops per evaluation, t_pass_ops per flip-pair pass it ran;
//   class key: per entry 20 per entry hash (2), 22 per symmetry (mix and 64-bit add).
static double expand_ops(const chz_ctx *c, int nq, const unsigned long long *hc) {
  const double E = (double)(1 << (2 * nq));
  const double child = E * (8.0 * c->avg_terms + 10.0), cheap = E * (4.0 + 5.0 * nq), cut = E * (30.0 + 3 * 14.0);
  // class key per lane: two 32-bit entry hashes (about 16 ops each), then per symmetry one slot summand (a 24-bit
  // and two 32-bit products with their shifts and xors, 13 ops) and its two-word share of the 64-bit sums
  const double key = E * (2 * 16.0 + 15.0 * c->ks.nsym);
  return (double)hc[C_EVAL] * child + (double)hc[C_CHEAP] * cheap + (double)hc[C_CUTS] * cut +
         (double)hc[C_NEEDT] * t_f0_ops(nq) + (double)hc[C_TPASS] * t_pass_ops(nq) + (double)hc[C_KEYS] * key;
}


constexpr uint32_t NO_PARENT = 0xffffffffu;
constexpr uint8_t CARRY_GATE = 0xfe;      // tree entry of a node carried unexpanded to the next level

// tree entries of n carried nodes: parent = their index in the level before (first_old + j)
__global__ void k_carry_tree(uint32_t *tp, uint8_t *tg, long long n, long long first_old) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x) {
    tp[j] = (uint32_t)(first_old + j);
    tg[j] = CARRY_GATE;
  }
}

// stable sort of (key prefix, rank) pairs of n survivors; the result lands in sk[0]/sv[0]
static void sort_pairs(chz_ctx *c, long long n) {
  const long long nb = (n + RS_TILE - 1) / RS_TILE;
  c->hist.need((size_t)nb * 256 + 1);
  c->scan_tmp.need((size_t)nb * 256 / SC_TILE + 4);
  HIPCHK(radix_sort_pairs(c->sk[0].p, c->sv[0].p, c->sk[1].p, c->sv[1].p, n, c->hist.p, c->scan_tmp.p, c->stream));
}

// bytes of the level control block for ns slots (its layout is the pinned staging area's)
static size_t ctl_in_bytes(int ns) { return ((size_t)C_N + (size_t)ns * IC_N) * 8 + (size_t)ns * 4 + (size_t)ns; }

// per-slot device arrays sized for `need` instance slots
static void reserve_slots(chz_ctx *c, int need) {
  if (need <= c->nslots) return;
  int ns = std::max(16, c->nslots);
  while (ns < need) ns *= 2;
  c->ctl.need((ctl_in_bytes(ns) + 7) / 8);
  c->cnt = c->ctl.p;
  c->icnt = c->cnt + C_N;
  c->d_rem = (int *)(c->icnt + (size_t)ns * IC_N);
  c->d_keepi = (uint8_t *)(c->d_rem + ns);
  c->d_active.need(ns);
  c->gflag.need(ns);
  HIPCHK(hipMemsetAsync(c->gflag.p, 0, (size_t)ns * 4, c->stream));   // fresh slots: no goal child known
  c->d_lo.need(ns);
  c->d_hi.need(ns);
  c->nslots = ns;
}

// nodes of every slot in frontier buffer `b` of m nodes (nodes are contiguous per instance)
__global__ void k_slot_bounds(const uint32_t *inst, long long n, long long *lo, long long *hi) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const uint32_t s = inst[i];
    if (i == 0 || inst[i - 1] != s) lo[s] = i;
    if (i == n - 1 || inst[i + 1] != s) hi[s] = i + 1;
  }
}
__global__ void k_fill_ll(long long *p, long long n, long long v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

static std::vector<long long> slot_counts(chz_ctx *c, const uint32_t *inst, long long m) {
  std::vector<long long> out(c->nslots);
  // (called after the level's control transfers completed: the staging area is free)
  long long *lo = (long long *)c->pin.need((size_t)c->nslots * 16), *hi = lo + c->nslots;
  hipLaunchKernelGGL(k_fill_ll, dim3(16), dim3(256), 0, c->stream, c->d_lo.p, (long long)c->nslots, 0ll);
  hipLaunchKernelGGL(k_fill_ll, dim3(16), dim3(256), 0, c->stream, c->d_hi.p, (long long)c->nslots, 0ll);
  if (m) hipLaunchKernelGGL(k_slot_bounds, dim3(grid_for(m / 64 + 1)), dim3(256), 0, c->stream, inst, m, c->d_lo.p, c->d_hi.p);
  HIPCHK(hipMemcpyAsync(lo, c->d_lo.p, c->nslots * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(hi, c->d_hi.p, c->nslots * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int s = 0; s < c->nslots; s++) out[s] = hi[s] - lo[s];
  return out;
}

// the next frontier: `distinct` gathered children (already in buffer nxt) followed by the admitted roots
template <int NQ> static void append_roots(chz_ctx *c, int nxt, long long distinct, uint32_t *tp, uint8_t *tg) {
  constexpr int E = 1 << (2 * NQ);
  const long long nr = (long long)c->pending.size();
  if (!nr) return;
  std::vector<int4> res(nr * E);
  std::vector<int> pk(nr);
  std::vector<uint8_t> none(nr, 255);
  std::vector<uint32_t> slot(nr), np(nr, NO_PARENT), bw(nr, BW_TUNK);
  for (long long i = 0; i < nr; i++) {
    const Pending &q = c->pending[i];
    for (int e = 0; e < E; e++) res[i * E + e] = q.U[e];
    pk[i] = q.k;
    slot[i] = (uint32_t)c->binst[q.inst].slot;
  }
  HIPCHK(hipMemcpyAsync(c->res[nxt].p + distinct * E, res.data(), nr * E * sizeof(int4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->pk[nxt].p + distinct, pk.data(), nr * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->g1[nxt].p + distinct, none.data(), nr, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->g2[nxt].p + distinct, none.data(), nr, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->finst[nxt].p + distinct, slot.data(), nr * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->pb[nxt].p + distinct, bw.data(), nr * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(tp + distinct, np.data(), nr * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(tg + distinct, none.data(), nr, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (long long i = 0; i < nr; i++) {
    c->binst[c->pending[i].inst].root_step = c->gstep + 1;
    c->forder.push_back(c->pending[i].inst);
  }
  c->pending.clear();
}

// the lexicographically first circuit of a FOUND child: gate of the child, then the tree walked up to the root
static void walk_circuit(chz_ctx *c, unsigned long long o, int G, std::vector<int> &seq) {
  seq.clear();
  seq.push_back((int)(o % G));
  uint32_t i = (uint32_t)(o / G);
  for (int s = c->gstep; s >= 0; s--) {
    uint32_t par;
    uint8_t g;
    chz::copy_sync(&par, c->tree_par[s] + i, 4, hipMemcpyDeviceToHost, c->stream);
    if (par == NO_PARENT) break;
    chz::copy_sync(&g, c->tree_gate[s] + i, 1, hipMemcpyDeviceToHost, c->stream);
    if (g != CARRY_GATE) seq.push_back(g);     // a level the search waited through
    i = par;
  }
  std::reverse(seq.begin(), seq.end());
}

// tree levels no live search can reach any more
static void prune_tree(chz_ctx *c) {
  int oldest = c->gstep;
  for (const BInst &b : c->binst) if (b.live) oldest = std::min(oldest, b.root_step);
  for (int s = 0; s < oldest && s < (int)c->tree_par.size(); s++) {
    if (c->tree_par[s]) { c->tree_free.push_back(c->tree_par[s]); c->tree_par[s] = nullptr; }
    c->tree_gate[s] = nullptr;
  }
}

// Captured level step for levels of at most STEP_CAP parents: expansion, survivor compaction, radix sort, exact
// dedupe, the searches' continue flags and the kept compaction, in one graph built once per frontier parity with
// launches sized for the capacity; every kernel reads the level's sizes from the counters the host uploads
// (C_M, C_NCC) and those the step produces (C_SURV, C_RUNS, the kept total).  The graph holds no memset node.  It
// is rebuilt when a buffer address, the slot count or the schedule it was captured with changes.  The gather runs
// after it as an ordinary launch (its tree level differs every step), then the level synchronises once.
constexpr long long STEP_CAP = 1024;
// Captures run while other threads (the request server's other engines) keep working on the device.  None of
// them may synchronise the whole device or allocate during a capture (a device-wide synchronisation includes the
// stream under capture; HIP's hipFree synchronises the device): buffer growth waits for the context's own stream
// only, and allocations, frees, device-wide synchronisations and memory queries take capture_mu, which the capture
// holds.
// Captured steps are used by batches that have held more than one search at once (the request server's engines):
// measured on MI355X, a single search runs its small levels faster as direct launches (on a reused context, Toffoli
// budget 9: 0.86 ms against 1.00 ms with captured steps, budget 12: 8.34 ms against 8.44 ms), while in the server
// they lower the latency tail.
constexpr bool kStepGraphs = true;
constexpr int STEP_BUCKET = 10;    // the schedule of the 1,024-parent bucket drives the captured expansion

template <int NQ>
static void step_graph(chz_ctx *c, int cur, long long m, const unsigned long long *hc_in, unsigned long long *hc,
                       unsigned long long *ic, unsigned long long *ptot, uint32_t *&tp, uint8_t *&tg) {
  constexpr int N = 1 << NQ, E = N * N, CPW = 64 / E;
  (void)N;
  const KSpec &ks = c->ks;
  const int G = ks.G, nxt = cur ^ 1, ns = c->nslots;
  const long long cap = STEP_CAP, ncap = cap * G;
  const long long nbA = (ncap + CT_TILE - 1) / CT_TILE, nbS = (ncap + RS_TILE - 1) / RS_TILE;
  const long long mx = ncap + (long long)c->pending.size();
  c->status.need(ncap); c->keys.need(ncap); c->info.need(ncap);
  c->bcount.need(nbA + 1);
  c->surv_o.need(ncap); c->surv_key.need(ncap); c->surv_bw.need(ncap);
  c->kept.need(ncap); c->kidx.need(ncap); c->runs.need(ncap);
  for (int b = 0; b < 2; b++) { c->sk[b].need(ncap); c->sv[b].need(ncap); }
  c->hist.need((size_t)nbS * 256 + 1);
  c->scan_tmp.need((size_t)nbS * 256 / SC_TILE + 4);
  c->res[nxt].need((size_t)mx * E); c->pk[nxt].need(mx); c->g1[nxt].need(mx); c->g2[nxt].need(mx);
  c->pb[nxt].need(mx); c->finst[nxt].need(mx);
  tree_alloc(c, mx, tp, tg);
  const chz_ctx::Sched sc = c->sched[STEP_BUCKET];
  const uint32_t *pinst = c->finst[cur].p;
  const std::vector<uintptr_t> key = {
      (uintptr_t)c->res[cur].p, (uintptr_t)c->pk[cur].p, (uintptr_t)c->g1[cur].p, (uintptr_t)c->g2[cur].p,
      (uintptr_t)c->pb[cur].p, (uintptr_t)pinst, (uintptr_t)c->status.p, (uintptr_t)c->keys.p, (uintptr_t)c->info.p,
      (uintptr_t)c->cnt, (uintptr_t)c->d_rem, (uintptr_t)c->d_keepi, (uintptr_t)c->icnt,
      (uintptr_t)c->bcount.p, (uintptr_t)c->surv_o.p, (uintptr_t)c->surv_key.p, (uintptr_t)c->surv_bw.p,
      (uintptr_t)c->kept.p, (uintptr_t)c->kidx.p, (uintptr_t)c->runs.p, (uintptr_t)c->sk[0].p, (uintptr_t)c->sk[1].p,
      (uintptr_t)c->sv[0].p, (uintptr_t)c->sv[1].p, (uintptr_t)c->hist.p, (uintptr_t)c->scan_tmp.p,
      (uintptr_t)c->d_active.p, (uintptr_t)c->gflag.p, (uintptr_t)ns, (uintptr_t)sc.var, (uintptr_t)sc.cap,
      (uintptr_t)sc.tw};
  chz_ctx::StepGraph &sg = c->sgraph[cur];
  if (!sg.exec || sg.key != key) {
    if (sg.exec) { HIPCHK(hipGraphExecDestroy(sg.exec)); sg.exec = nullptr; }
    const unsigned long long *dm = c->cnt + C_M, *dncc = c->cnt + C_NCC, *dsurv = c->cnt + C_SURV;
    hipStream_t s = c->stream;
    hipGraph_t graph = nullptr;
    // no device-wide or allocating HIP call from another thread while this stream is captured (capture_mu)
    std::lock_guard<std::recursive_mutex> capture_lock(capture_mu());
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    ExpandArgs ea{c->res[cur].p, c->pk[cur].p, c->g1[cur].p, c->g2[cur].p, 0, ncap, 0, 0,
                  c->status.p, c->keys.p, c->info.p, c->pb[cur].p, c->cnt, pinst, c->d_rem, c->d_keepi,
                  c->icnt, 1, dm, sc.tw, c->goal_skip ? c->gflag.p : nullptr};
    const dim3 grid(grid_for((cap * G + CPW - 1) / CPW, sc.cap));
    switch (sc.var) {
      case 0: hipLaunchKernelGGL((k_level<NQ, 4>), grid, dim3(256), 0, s, ks, ea, cap); break;
      case 2: hipLaunchKernelGGL((k_level<NQ, 6>), grid, dim3(256), 0, s, ks, ea, cap); break;
      default: hipLaunchKernelGGL((k_level<NQ, 5>), grid, dim3(256), 0, s, ks, ea, cap); break;
    }
    const FlagAlive fa{c->status.p, pinst, c->d_keepi, 0, G};
    hipLaunchKernelGGL(k_flag_count<FlagAlive>, dim3((unsigned)nbA), dim3(CT_THREADS), 0, s, fa, ncap, c->bcount.p, dncc);
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, s, c->bcount.p, nbA);
    hipLaunchKernelGGL(k_scatter_alive, dim3((unsigned)nbA), dim3(CT_THREADS), 0, s, fa, ncap, c->bcount.p, c->keys.p,
                       0ll, c->cnt + C_SURV, c->surv_o.p, c->surv_key.p, c->info.p, c->surv_bw.p, dncc);
    hipLaunchKernelGGL(k_add_total, dim3(1), dim3(64), 0, s, c->cnt + C_SURV, c->bcount.p + nbA);
    hipLaunchKernelGGL(k_init_sort, dim3(grid_for(ncap / 64 + 1)), dim3(256), 0, s, c->surv_key.p, ncap, c->sk[0].p,
                       c->sv[0].p, dsurv);
    hipError_t e = radix_sort_pairs(c->sk[0].p, c->sv[0].p, c->sk[1].p, c->sv[1].p, ncap, c->hist.p, c->scan_tmp.p, s,
                                    dsurv);
    hipLaunchKernelGGL(k_fill_kept, dim3(grid_for(ncap / 256 + 1)), dim3(256), 0, s, c->kept.p, ncap, dsurv, nullptr,
                       0);
    hipLaunchKernelGGL(k_run_starts, dim3(grid_for(ncap / (64 * RH_ITEMS) + 1)), dim3(256), 0, s, c->sk[0].p, ncap,
                       c->runs.p, c->cnt + C_RUNS, dsurv);
    const DedupeArgs da{c->sk[0].p, c->sv[0].p, ncap, c->surv_o.p, c->surv_key.p, c->runs.p, c->res[cur].p,
                        c->pk[cur].p, c->kept.p, c->cnt, pinst, dsurv};
    hipLaunchKernelGGL(k_dedupe_runs<NQ>, dim3(grid_for(ncap / CPW + 1, 4096)), dim3(256), 0, s, ks, da);
    hipLaunchKernelGGL(k_active, dim3(1), dim3(256), 0, s, c->icnt, c->d_keepi, ns, c->d_active.p);
    const FlagKept fk{c->kept.p, c->surv_o.p, pinst, c->d_active.p, G};
    hipLaunchKernelGGL(k_flag_count<FlagKept>, dim3((unsigned)nbA), dim3(CT_THREADS), 0, s, fk, ncap, c->bcount.p, dsurv);
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, s, c->bcount.p, nbA);
    hipLaunchKernelGGL(k_scatter_idx<FlagKept>, dim3((unsigned)nbA), dim3(CT_THREADS), 0, s, fk, ncap, c->bcount.p,
                       c->kidx.p, dsurv);
    if (e == hipSuccess) e = hipGetLastError();
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (e == hipSuccess) e = ec;
    if (e == hipSuccess) e = hipGraphInstantiate(&sg.exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      sg.exec = nullptr;
      set_err(std::string("HIP error: ") + hipGetErrorString(e) + " while capturing a level step");
      throw DeviceError();
    }
    sg.key = key;
  }
  HIPCHK(hipMemcpyAsync(c->cnt, hc_in, C_N * 8, hipMemcpyHostToDevice, c->stream));
  c->ev.begin(K_STEP, c->stream);
  HIPCHK(hipGraphLaunch(sg.exec, c->stream));
  c->ev.end(c->stream);
  const GatherArgs ga{c->kidx.p, ncap, c->surv_o.p, c->res[cur].p, c->pk[cur].p, c->g1[cur].p,
                      c->res[nxt].p, c->pk[nxt].p, c->g1[nxt].p, c->g2[nxt].p, tp, tg, pinst, c->finst[nxt].p,
                      c->surv_bw.p, c->pb[nxt].p, c->bcount.p + nbA, c->goal_skip ? c->gflag.p : nullptr, CPW};
  HIPCHK(hipMemsetAsync(c->gflag.p, 0, (size_t)ns * 4, c->stream));   // the graph's expansion has read them
  c->ev.begin(K_GATHER, c->stream);
  hipLaunchKernelGGL(k_gather<NQ>, dim3(grid_for((ncap + CPW - 1) / CPW)), dim3(256), 0, c->stream, ks, ga);
  c->ev.end(c->stream);
  HIPCHK(hipMemcpyAsync(hc, c->cnt, C_N * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(ic, c->icnt, (size_t)ns * IC_N * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(ptot, c->bcount.p + nbA, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->ev.flush(c->k_ms, c->k_launch);
  if (hc[C_REPOVF]) { set_err("a dedupe run holds more distinct classes than its buffer"); throw DeviceError(); }
  c->k_work[K_GATHER] += (double)*ptot * (E * 16.0 * 2.0 + 16.0);
}

// One level of every live search of the batch (src/cheese.cu:1249-1377 per search): expansion in chunks,
// survivor compaction, sort, exact dedupe, next frontier; then the searches admitted since the last step join
// the new frontier as roots.  Level lines and results go to the context's event queue.
template <int NQ> static void batch_step(chz_ctx *c) {
  constexpr int N = 1 << NQ, E = N * N, CPW = 64 / E;
  const KSpec &ks = c->ks;
  const int G = ks.G;
  if (c->m == 0 && c->pending.empty()) return;
  if (c->m == 0) {   // a frontier of roots only
    if (c->nslots) HIPCHK(hipMemsetAsync(c->gflag.p, 0, (size_t)c->nslots * 4, c->stream));
    const long long nr = (long long)c->pending.size();
    const int nxt = c->cur ^ 1;
    c->res[nxt].need(nr * E); c->pk[nxt].need(nr); c->g1[nxt].need(nr); c->g2[nxt].need(nr); c->finst[nxt].need(nr);
    c->pb[nxt].need(nr);
    uint32_t *tp = nullptr;
    uint8_t *tg = nullptr;
    tree_alloc(c, nr, tp, tg);
    c->tree_par.push_back(tp);
    c->tree_gate.push_back(tg);
    c->gstep = (int)c->tree_par.size() - 1;
    c->gstep--;                 // append_roots marks roots with gstep + 1
    append_roots<NQ>(c, nxt, 0, tp, tg);
    c->gstep++;
    c->cur = nxt;
    c->m = nr;
    const std::vector<long long> cnts = slot_counts(c, c->finst[c->cur].p, c->m);
    for (BInst &b : c->binst) if (b.live) b.nodes = cnts[b.slot];
  }
  const double tl = now_s();
  const int cur = c->cur;
  const long long m = c->m;
  // per-level memory planning: the searches expanded this step are the longest prefix of the frontier (whole
  // searches, in frontier order, at least one) whose planned bytes fit the budget beside the frontier held: per
  // parent its children's records and its share of the next frontier, by the search's last growth (a prior of
  // kGrowthPrior before its first level); the others wait a step, carried unexpanded
  long long ma = m;
  std::vector<char> carried(c->binst.size(), 0);
  if (c->mem_budget > 0 && m > 0) {
    constexpr double node_b = E * 16.0 + 24.0, child_b = 48.0, kGrowthPrior = 8.0;
    const double held = (double)m * node_b;
    double plan = 0;
    long long acc = 0;
    bool stop = false;
    for (int bi : c->forder) {
      const BInst &b = c->binst[bi];
      if (!b.live) continue;
      const double g = std::min<double>(G, b.growth > 0 ? b.growth : kGrowthPrior);
      const double need = (double)b.nodes * (G * child_b + g * node_b);
      if (acc > 0 && (stop || held + plan + need > c->mem_budget)) { stop = true; carried[bi] = 1; continue; }
      plan += need;
      acc += b.nodes;
    }
    ma = acc;
  }
  // the level's control block staged in pinned memory with the device block's layout, uploaded in one copy:
  // [counters in | per-slot counters in | rem | keepi], then [counters out | per-slot counters out] (read back
  // in one copy), the kept total and the per-slot active flags
  const int ns = c->nslots;
  const size_t in_bytes = ctl_in_bytes(ns), o_out = (in_bytes + 63) & ~(size_t)63,
               o_tot = o_out + ((size_t)C_N + (size_t)ns * IC_N) * 8, o_act = o_tot + 8;
  uint8_t *pin = c->pin.need(o_act + ns);
  unsigned long long *hc_in = (unsigned long long *)pin, *ic_in = hc_in + C_N;
  unsigned long long *hc = (unsigned long long *)(pin + o_out), *ic = hc + C_N;
  unsigned long long *ptot = (unsigned long long *)(pin + o_tot);
  int *rem = (int *)(ic_in + (size_t)ns * IC_N);
  uint8_t *keepi = (uint8_t *)(rem + ns), *active = pin + o_act;
  for (int i = 0; i < C_N; i++) hc_in[i] = 0;
  for (int s = 0; s < ns; s++) {
    rem[s] = 0;
    keepi[s] = active[s] = 0;
    for (int i = 0; i < IC_N; i++) ic_in[(size_t)s * IC_N + i] = i == IC_FIRST ? ~0ull : 0ull;
  }
  for (const BInst &b : c->binst)
    if (b.live) { rem[b.slot] = b.depth - (b.t + 1); keepi[b.slot] = b.t + 1 < b.depth; }
  HIPCHK(hipMemcpyAsync(c->ctl.p, pin, in_bytes, hipMemcpyHostToDevice, c->stream));
  const long long chunk = c->opts.chunk > 0 ? c->opts.chunk : (1ll << 20);
  const uint32_t *pinst = c->finst[cur].p;
  const int nxt = cur ^ 1;
  long long nsurv = 0, distinct = 0;
  bool gflag_clear = false;          // the goal flags were cleared by this level's kept fill
  uint32_t *tp = nullptr;
  uint8_t *tg = nullptr;
  // small levels run as one captured graph (expansion through kept compaction) plus the gather: one
  // synchronisation per level instead of two, and one launch instead of about twenty
  const bool graphed = kStepGraphs && c->shared_steps && !c->tuning && !c->gt_filter && c->opts.dedupe && m > 0 &&
                       m <= STEP_CAP && m <= chunk && ma == m;
  if (graphed) {
    for (int i = 0; i < C_N; i++) hc_in[i] = 0;
    hc_in[C_M] = (unsigned long long)m;
    hc_in[C_NCC] = (unsigned long long)(m * G);
    step_graph<NQ>(c, cur, m, hc_in, hc, ic, ptot, tp, tg);
    nsurv = (long long)hc[C_SURV];
    distinct = (long long)*ptot;
    c->k_work[K_STEP] += expand_ops(c, NQ, hc) + (double)(m * G) + (double)nsurv * (24.0 + 16.0 + 4 * 20.0 + 9.0);
    static const int stage_ctr[6] = {C_EVAL, C_CHEAP, C_CUTS, C_NEEDT, C_KEYS, C_TPASS};
    for (int i = 0; i < 6; i++) c->k_stage[i] += (double)hc[stage_ctr[i]];
    for (const BInst &b : c->binst)
      if (b.live) {
        const unsigned long long *x = &ic[(size_t)b.slot * IC_N];
        active[b.slot] = !x[IC_FOUND] && !x[IC_OVF] && keepi[b.slot];
      }
  } else {
  const long long maxcc = std::min<long long>(chunk, std::max<long long>(ma, 1)) * G;
  c->status.need(maxcc);
  c->info.need(maxcc);
  c->keys.need(maxcc);
  c->bcount.need((maxcc + CT_TILE - 1) / CT_TILE + 1);
  // survivors of every chunk append at the device-side count C_SURV: the buffers hold every child of the level
  c->surv_o.need((size_t)std::max<long long>(ma, 1) * G);
  c->surv_key.need((size_t)std::max<long long>(ma, 1) * G);
  c->surv_bw.need((size_t)std::max<long long>(ma, 1) * G);
  for (long long off = 0; off < ma; off += chunk) {
    const long long cm = std::min(chunk, ma - off), ncc = cm * G;
    // the schedule of this chunk's size bucket: few parents split each parent's candidates over several lane
    // groups so the level fills the GPU
    const int bucket = std::min(31, 63 - __builtin_clzll((unsigned long long)cm));
    const chz_ctx::Sched sc = c->tuning ? c->force : c->sched[bucket];
    const int gsplit = (int)std::max<long long>(1, std::min<long long>(G, ((long long)sc.tw * CPW + cm - 1) / cm));
    ExpandArgs ea{c->res[cur].p, c->pk[cur].p, c->g1[cur].p, c->g2[cur].p, off, ncc, 0, 0,
                  c->status.p, c->keys.p, c->info.p, c->pb[cur].p, c->cnt, pinst, c->d_rem, c->d_keepi,
                  c->icnt, gsplit, nullptr, 0, c->goal_skip ? c->gflag.p : nullptr};
    const dim3 grid(grid_for((cm * gsplit + CPW - 1) / CPW, sc.cap));
    hipEvent_t te0 = nullptr, te1 = nullptr;
    if (c->tuning) {
      te0 = c->ev.take();
      te1 = c->ev.take();
      HIPCHK(hipEventRecord(te0, c->stream));
    }
    c->ev.begin(K_EXPAND, c->stream);
    switch (sc.var) {
      case 0: hipLaunchKernelGGL((k_level<NQ, 4>), grid, dim3(256), 0, c->stream, ks, ea, cm); break;
      case 2: hipLaunchKernelGGL((k_level<NQ, 6>), grid, dim3(256), 0, c->stream, ks, ea, cm); break;
      default: hipLaunchKernelGGL((k_level<NQ, 5>), grid, dim3(256), 0, c->stream, ks, ea, cm); break;
    }
    c->ev.end(c->stream);
    if (c->tuning) {
      HIPCHK(hipEventRecord(te1, c->stream));
      c->tune_ev.push_back({bucket, {te0, te1}});
    }
    const long long nb = (ncc + CT_TILE - 1) / CT_TILE;
    const FlagAlive fa{c->status.p, pinst, c->d_keepi, off, G};
    c->ev.begin(K_COMPACT, c->stream);
    if (nb == 1) {   // one tile: count, place and advance the survivor count in one launch
      hipLaunchKernelGGL(k_scatter_alive_one, dim3(1), dim3(CT_THREADS), 0, c->stream, fa, ncc, c->keys.p, off * G,
                         c->cnt + C_SURV, c->surv_o.p, c->surv_key.p, c->info.p, c->surv_bw.p);
    } else {
      hipLaunchKernelGGL(k_flag_count<FlagAlive>, dim3((unsigned)nb), dim3(CT_THREADS), 0, c->stream, fa, ncc,
                         c->bcount.p);
      hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, c->stream, c->bcount.p, nb);
      hipLaunchKernelGGL(k_scatter_alive, dim3((unsigned)nb), dim3(CT_THREADS), 0, c->stream, fa, ncc, c->bcount.p,
                         c->keys.p, off * G, c->cnt + C_SURV, c->surv_o.p, c->surv_key.p, c->info.p, c->surv_bw.p,
                         nullptr);
      hipLaunchKernelGGL(k_add_total, dim3(1), dim3(64), 0, c->stream, c->cnt + C_SURV, c->bcount.p + nb);
    }
    c->ev.end(c->stream);
    c->k_work[K_COMPACT] += (double)ncc;
  }
  HIPCHK(hipMemcpyAsync(hc, c->cnt, ((size_t)C_N + (size_t)ns * IC_N) * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->ev.flush(c->k_ms, c->k_launch);
  nsurv = (long long)hc[C_SURV];
  c->k_work[K_COMPACT] += (double)nsurv * 24.0;
  c->k_work[K_EXPAND] += expand_ops(c, NQ, hc);
  static const int stage_ctr[6] = {C_EVAL, C_CHEAP, C_CUTS, C_NEEDT, C_KEYS, C_TPASS};
  for (int i = 0; i < 6; i++) c->k_stage[i] += (double)hc[stage_ctr[i]];
  // searches that end at this level with a circuit or an overflow; the rest continue if their budget allows
  for (const BInst &b : c->binst)
    if (b.live) {
      const unsigned long long *x = &ic[(size_t)b.slot * IC_N];
      active[b.slot] = !x[IC_FOUND] && !x[IC_OVF] && keepi[b.slot];
    }
  HIPCHK(hipMemcpyAsync(c->d_active.p, active, ns, hipMemcpyHostToDevice, c->stream));
  if (nsurv > 0) {
    c->kept.need(nsurv);
    c->kidx.need(nsurv);
    if (c->opts.dedupe && nsurv > 1) {
      for (int bb = 0; bb < 2; bb++) { c->sk[bb].need(nsurv); c->sv[bb].need(nsurv); }
      c->runs.need(nsurv);
      c->ev.begin(K_SORT, c->stream);
      if (nsurv <= SS_MAX) {
        hipLaunchKernelGGL(k_sort_small, dim3(1), dim3(1024), 0, c->stream, c->surv_key.p, nsurv, c->sk[0].p,
                           c->sv[0].p);
      } else {
        hipLaunchKernelGGL(k_init_sort, dim3(grid_for(nsurv / 64 + 1)), dim3(256), 0, c->stream, c->surv_key.p,
                           nsurv, c->sk[0].p, c->sv[0].p, nullptr);
        sort_pairs(c, nsurv);
      }
      c->ev.end(c->stream);
      c->k_work[K_SORT] += (double)nsurv * (16.0 + 4 * (4.0 + 16.0));
      c->ev.begin(K_DEDUPE, c->stream);
      hipLaunchKernelGGL(k_fill_kept, dim3(grid_for(nsurv / 256 + 1)), dim3(256), 0, c->stream, c->kept.p, nsurv,
                         nullptr, c->gflag.p, ns);
      gflag_clear = true;
      hipLaunchKernelGGL(k_run_starts, dim3(grid_for(nsurv / (64 * RH_ITEMS) + 1)), dim3(256), 0, c->stream,
                         c->sk[0].p, nsurv, c->runs.p, c->cnt + C_RUNS, nullptr);
      const DedupeArgs da{c->sk[0].p, c->sv[0].p, nsurv, c->surv_o.p, c->surv_key.p, c->runs.p,
                          c->res[cur].p, c->pk[cur].p, c->kept.p, c->cnt, pinst, nullptr};
      hipLaunchKernelGGL(k_dedupe_runs<NQ>, dim3(grid_for(nsurv / CPW + 1, 4096)), dim3(256), 0, c->stream, ks, da);
      c->ev.end(c->stream);
      c->k_work[K_DEDUPE] += (double)nsurv * 9.0;
    } else {
      HIPCHK(hipMemsetAsync(c->kept.p, 1, nsurv, c->stream));
    }
    if (c->gt_filter)      // a table build: classes already in the table (lower levels) are dropped
      hipLaunchKernelGGL(k_gt_filter<NQ>, dim3(grid_for(nsurv / CPW + 1)), dim3(256), 0, c->stream, ks,
                         c->gt_filter->dev(), c->surv_o.p, c->surv_key.p, nsurv, c->res[cur].p, c->pk[cur].p,
                         c->kept.p);
    const long long nb = (nsurv + CT_TILE - 1) / CT_TILE;
    c->bcount.need(nb + 1);
    const FlagKept fk{c->kept.p, c->surv_o.p, pinst, c->d_active.p, G};
    c->ev.begin(K_COMPACT, c->stream);
    if (nb == 1) {
      hipLaunchKernelGGL(k_compact_one<FlagKept>, dim3(1), dim3(CT_THREADS), 0, c->stream, fk, nsurv, c->kidx.p,
                         c->bcount.p + nb);
    } else {
      hipLaunchKernelGGL(k_flag_count<FlagKept>, dim3((unsigned)nb), dim3(CT_THREADS), 0, c->stream, fk, nsurv,
                         c->bcount.p);
      hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, c->stream, c->bcount.p, nb);
      hipLaunchKernelGGL(k_scatter_idx<FlagKept>, dim3((unsigned)nb), dim3(CT_THREADS), 0, c->stream, fk, nsurv,
                         c->bcount.p, c->kidx.p);
    }
    c->ev.end(c->stream);
    HIPCHK(hipMemcpyAsync(ptot, c->bcount.p + nb, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hc, c->cnt, C_N * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (hc[C_REPOVF]) { set_err("a dedupe run holds more distinct classes than its buffer"); throw DeviceError(); }
    distinct = (long long)*ptot;
  }
  }   // normal path
  // next frontier: kept children, then the searches admitted since the last step.  The goal flags describe the
  // frontier being expanded until here; the gather sets them for the next one.
  if (!graphed && !gflag_clear) HIPCHK(hipMemsetAsync(c->gflag.p, 0, (size_t)ns * 4, c->stream));
  const long long ncarry = m - ma;            // nodes of the searches that waited
  const long long mn = distinct + ncarry + (long long)c->pending.size();
  std::vector<long long> cnts(c->nslots, 0);
  if (mn) {
    if (!graphed) {
      c->res[nxt].need((size_t)mn * E); c->pk[nxt].need(mn); c->g1[nxt].need(mn); c->g2[nxt].need(mn);
      c->pb[nxt].need(mn);
      c->finst[nxt].need(mn);
      tree_alloc(c, mn, tp, tg);
    }
    if (distinct && !graphed) {
      const GatherArgs ga{c->kidx.p, distinct, c->surv_o.p, c->res[cur].p, c->pk[cur].p, c->g1[cur].p,
                          c->res[nxt].p, c->pk[nxt].p, c->g1[nxt].p, c->g2[nxt].p, tp, tg, pinst, c->finst[nxt].p,
                          c->surv_bw.p, c->pb[nxt].p, nullptr, c->goal_skip ? c->gflag.p : nullptr,
                          gather_tile<NQ>(distinct)};
      c->ev.begin(K_GATHER, c->stream);
      hipLaunchKernelGGL(k_gather<NQ>, dim3(grid_for((distinct + ga.tn - 1) / ga.tn)), dim3(256), 0, c->stream, ks,
                         ga);
      c->ev.end(c->stream);
      c->k_work[K_GATHER] += (double)distinct * (E * 16.0 * 2.0 + 16.0);
    }
    const bool single = c->pending.empty() && chz_batch_live(c) == 1;
    if (c->gt_filter && distinct)   // a table build keeps every level, resident, with its classes' keys
      gt_add_level<NQ>(ks, *c->gt_filter, c->res[nxt].p, c->pk[nxt].p, tp, tg, distinct, c->stream);
    if (ncarry) {                   // the waiting searches' nodes follow the gathered ones, unchanged
      HIPCHK(hipMemcpyAsync(c->res[nxt].p + (size_t)distinct * E, c->res[cur].p + (size_t)ma * E,
                            (size_t)ncarry * E * sizeof(int4), hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->pk[nxt].p + distinct, c->pk[cur].p + ma, (size_t)ncarry * 4, hipMemcpyDeviceToDevice,
                            c->stream));
      HIPCHK(hipMemcpyAsync(c->g1[nxt].p + distinct, c->g1[cur].p + ma, (size_t)ncarry, hipMemcpyDeviceToDevice,
                            c->stream));
      HIPCHK(hipMemcpyAsync(c->g2[nxt].p + distinct, c->g2[cur].p + ma, (size_t)ncarry, hipMemcpyDeviceToDevice,
                            c->stream));
      HIPCHK(hipMemcpyAsync(c->pb[nxt].p + distinct, c->pb[cur].p + ma, (size_t)ncarry * 4, hipMemcpyDeviceToDevice,
                            c->stream));
      HIPCHK(hipMemcpyAsync(c->finst[nxt].p + distinct, c->finst[cur].p + ma, (size_t)ncarry * 4,
                            hipMemcpyDeviceToDevice, c->stream));
      hipLaunchKernelGGL(k_carry_tree, dim3(grid_for(ncarry / 64 + 1)), dim3(256), 0, c->stream, tp + distinct,
                         tg + distinct, ncarry, ma);
      for (size_t bi = 0; bi < carried.size(); bi++) c->carried_levels += carried[bi];
    }
    append_roots<NQ>(c, nxt, distinct + ncarry, tp, tg);
    c->tree_par.push_back(tp);
    c->tree_gate.push_back(tg);
    if (single) {   // one search: its nodes are the whole frontier (its kernel times are read next step)
      for (const BInst &b : c->binst) if (b.live) cnts[b.slot] = mn;
    } else {
      cnts = slot_counts(c, c->finst[nxt].p, mn);
      c->ev.flush(c->k_ms, c->k_launch);
    }
  } else {
    if (graphed && tp) c->tree_free.push_back(tp);     // the level's tree block holds no node
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ev.flush(c->k_ms, c->k_launch);
  }
  // events of this level, per search
  const double now = now_s();
  for (size_t bi = 0; bi < c->binst.size(); bi++) {
    BInst &b = c->binst[bi];
    if (!b.live || b.root_step > c->gstep) continue;      // admitted in this step: it has not expanded yet
    if (bi < carried.size() && carried[bi]) continue;      // waited this step: no level
    const unsigned long long *x = &ic[(size_t)b.slot * IC_N];
    chz_event evn{};
    evn.tag = b.tag;
    if (x[IC_OVF]) {
      evn.kind = 1;
      evn.result.rc = 4;
      evn.result.seconds = now - b.t0;
      set_err("residual coefficients exceed the storage type (int32)");
    } else if (x[IC_FOUND]) {
      std::vector<int> seq;
      walk_circuit(c, x[IC_FIRST], G, seq);
      evn.kind = 1;
      evn.result.rc = 0;
      evn.result.len = (int)seq.size();
      evn.result.found = (long)x[IC_FOUND];
      for (size_t j = 0; j < seq.size() && j < CHZ_DMAX; j++) evn.result.circuit[j] = seq[j];
      evn.result.seconds = now - b.t0;
    } else {
      const long long dist = active[b.slot] ? cnts[b.slot] : 0;
      chz_event lv{};
      lv.tag = b.tag;
      lv.kind = 0;
      lv.level.t = b.t + 1;
      lv.level.parents = (long)b.nodes;
      lv.level.prefixes = (long)(b.nodes * G - (long long)x[IC_W2]);
      lv.level.cut_w3 = (long)x[IC_W3];
      lv.level.cut_lb = (long)x[IC_DEAD];
      lv.level.survivors = (long)x[IC_ALIVE];
      lv.level.distinct = (long)dist;
      lv.level.proved_unsat = dist ? 0 : 1;
      lv.level.seconds = now - tl;
      c->events.push_back(lv);
      b.t++;
      if (b.nodes > 0) b.growth = (double)dist / (double)b.nodes;
      b.nodes = dist;
      if (dist) continue;
      evn.kind = 1;
      evn.result.rc = 1;
      evn.result.seconds = now - b.t0;
    }
    c->events.push_back(evn);
    b.live = false;
  }
  // roots admitted in this step
  for (BInst &b : c->binst)
    if (b.live && b.root_step == c->gstep + 1) b.nodes = cnts[b.slot];
  c->cur = nxt;
  c->m = mn;
  c->gstep++;
  c->forder.erase(std::remove_if(c->forder.begin(), c->forder.end(), [&](int bi) { return !c->binst[bi].live; }),
                  c->forder.end());
  prune_tree(c);
}

// LevelTabs for a spec: the term descriptors and headers of its candidates, its symmetries' entry positions (n = 3),
// its cx table (when it fits) and the entry sources of its cut matrices
static void build_level_tabs(LevelTabs &T, const std::vector<DevGate> &dg, const KSpec &k, const chz::SymGroup &grp,
                             const HostSpec &s) {
  std::memset(&T, 0, sizeof T);
  const int n = k.n, N = 1 << n, GS = N * N;
  for (int g = 0; g < k.G; g++) {
    const DevGate &d = dg[g];
    for (int e = 0; e < GS; e++) {
      const int ec = e & (N - 1);
      int lc = (ec >> d.q[0]) & 1;
      if (d.nq > 1) lc |= ((ec >> d.q[1]) & 1) << 1;
      if (d.nq > 2) lc |= ((ec >> d.q[2]) & 1) << 2;
      const int cb = ec & ~d.mask;
      uint32_t desc = 0;
      int nt = 0;
      for (int m = 0; d.unit && m < d.L; m++) {
        const int p = d.ph[m * d.L + lc];
        if (p < 0) continue;
        int x = cb | ((m & 1) << d.q[0]);
        if (d.nq > 1) x |= ((m >> 1) & 1) << d.q[1];
        if (d.nq > 2) x |= ((m >> 2) & 1) << d.q[2];
        if (nt < 2) desc |= (uint32_t)(x | (p << 3)) << (6 * nt);
        nt++;
      }
      T.sdesc[g * 64 + e] = (uint16_t)(!d.unit || nt == 0 || nt > 2 ? 0x8000u : nt == 2 ? desc | 0x1000u : desc);
    }
    int kind = 0;
    if (d.unit) {
      kind = 2;                                // permutation: one nonzero per column
      for (int lc = 0; lc < d.L; lc++) {
        int nz = 0;
        for (int m = 0; m < d.L; m++) nz += d.ph[m * d.L + lc] >= 0;
        if (nz != 1) kind = 1;
      }
    }
    T.shdr[g] = GateHdr{d.nq, d.L, d.mask, d.ik, d.q[0], d.q[1], d.q[2], kind, d.tdel, d.cutloc, d.cdel, d.cls};
  }
  if (n == 3)
    for (int y = 0; y < 16 && y < k.nsym; y++)
      for (int e = 0; e < 64; e++) T.spos[y * 64 + e] = (uint8_t)(grp.bperm[y * 8 + (e >> 3)] * 8 + grp.bperm[y * 8 + (e & 7)]);
  if (k.cxlds && k.have_table)
    for (int i = 0; i < CXL_MAX && i < k.ncx; i++) T.scx[i] = (uint8_t)s.cxtable[i];
  for (int i = 0; i < CSRC_MAX * GS; i++) {
    if (i / GS >= k.ncuts) continue;
    const int le = i % GS, A = k.cutmask[i / GS], sa = k.cutsize[i / GS];
    T.scsrc[i] = (uint8_t)(n == 3 ? cut_source<3>(le, A, sa) : n == 2 ? cut_source<2>(le, A, sa) : cut_source<1>(le, A, sa));
  }
}

static void upload_spec(chz_ctx *c) {
  const HostSpec &s = c->hs;
  const int N = s.N(), G = s.G();
  std::vector<DevGate> dg(G);
  double terms = 0, columns = 0;
  for (int g = 0; g < G; g++) {
    const HostGate &h = s.gates[g];
    DevGate &d = dg[g];
    std::memset(&d, 0, sizeof d);
    d.nq = h.nq;
    d.L = 1 << h.nq;
    d.ik = h.ik;
    for (int j = 0; j < h.nq; j++) { d.q[j] = h.q[j]; d.mask |= 1 << h.q[j]; }
    d.unit = 1;
    for (int i = 0; i < d.L * d.L; i++) {
      const Zq &z = h.inv[i];
      d.inv[i] = make_int4((int)z.a, (int)z.b, (int)z.c, (int)z.d);
      // +-w^p has exactly one coefficient equal to +-1: w^p for p < 4 at index p, -w^p = w^(p+4)
      const int nz = (z.a != 0) + (z.b != 0) + (z.c != 0) + (z.d != 0);
      if (!nz) { d.ph[i] = -1; continue; }
      const int64_t cf[4] = {z.a, z.b, z.c, z.d};
      int idx = -1;
      for (int j = 0; j < 4; j++) if (cf[j]) idx = j;
      if (nz != 1 || (cf[idx] != 1 && cf[idx] != -1)) { d.unit = 0; d.ph[i] = -1; continue; }
      d.ph[i] = (int8_t)(idx + (cf[idx] < 0 ? 4 : 0));
    }
    // a Clifford gate's Pauli transfer matrix is a signed permutation, so a child's T bound (the smallest
    // sqrt2-valuation over the transfer matrix of the residual) is its parent's; a gate on one side of a cut
    // leaves that cut's operator-Schmidt rank unchanged.  Packed cut needs take 3 bits each in the node's
    // bound word, so at most BW_MAXCUTS cuts are carried.
    const int sde = gate_ptm_sde(h);
    d.tdel = sde >= 0 && sde < 255 ? sde : 255;
    for (const Zq &z : h.inv) terms += (z.a | z.b | z.c | z.d) != 0;
    columns += d.L;
    d.cutloc = 0;
    d.cdel = 0;
    if ((int)s.cutmask.size() <= BW_MAXCUTS)
      for (size_t i = 0; i < s.cutmask.size(); i++) {
        if ((d.mask & s.cutmask[i]) == 0 || (d.mask & ~s.cutmask[i]) == 0) d.cutloc |= 1 << i;
        const int sr = gate_cut_rank(h, s.cutmask[i]);
        int del = 0;
        while (del < 7 && (1 << del) < sr) del++;
        d.cdel |= (sr < 1 ? 7 : del) << (3 * i);
      }
  }
  c->avg_terms = columns > 0 ? terms / columns : 1.0;
  // cut-delta classes: the lower cut bound of a child depends on its parent and on its gate's cdel only, so the
  // expansion computes it once per parent and class
  std::vector<uint32_t> cdv;
  for (DevGate &d : dg) {
    size_t j = 0;
    while (j < cdv.size() && cdv[j] != (uint32_t)d.cdel) j++;
    if (j == cdv.size()) cdv.push_back((uint32_t)d.cdel);
    d.cls = (int)j;
  }
  c->cdv = cdv;
  // T-bound classes: non-Clifford candidates whose children always have equal T bounds (same_t_bound), so a
  // parent's expansion computes the bound once per class
  {
    std::vector<int> cls(G, -1);
    int ncls = 0;
    for (int g = 0; g < G; g++) {
      if (dg[g].tdel == 0 || cls[g] >= 0) continue;
      std::vector<int> mem{g};
      for (int h2 = g + 1; h2 < G; h2++)
        if (dg[h2].tdel != 0 && cls[h2] < 0 && same_t_bound(s.gates[g], s.gates[h2])) mem.push_back(h2);
      if (mem.size() < 2 || ncls >= TC_MAX) continue;
      for (int m : mem) cls[m] = ncls;
      ncls++;
    }
    for (int g = 0; g < G; g++) dg[g].cls |= (cls[g] + 1) << 8;
  }
  std::vector<int4> goal(N * N);
  for (int i = 0; i < N * N; i++) goal[i] = make_int4((int)s.goal[i].a, (int)s.goal[i].b, (int)s.goal[i].c, (int)s.goal[i].d);
  std::vector<int> cx = s.cxtable.empty() ? std::vector<int>{0} : s.cxtable;
  // every table in one device block (256-byte aligned sections), one copy
  std::vector<uint8_t> blob;
  std::vector<std::pair<void **, size_t>> at;
  auto add = [&](void **dst, const void *src, size_t bytes) {
    const size_t off = (blob.size() + 255) & ~(size_t)255;
    blob.resize(off + std::max<size_t>(bytes, 1), 0);
    if (bytes) std::memcpy(blob.data() + off, src, bytes);
    at.push_back({dst, off});
  };
  add((void **)&c->d_gates, dg.data(), dg.size() * sizeof(DevGate));
  add((void **)&c->d_bad2, s.bad2.data(), s.bad2.size());
  add((void **)&c->d_bad3, s.bad3.data(), s.bad3.size() * 4);
  add((void **)&c->d_goal, goal.data(), goal.size() * sizeof(int4));
  add((void **)&c->d_cx, cx.data(), cx.size() * 4);
  add((void **)&c->d_symperm, c->grp.bperm.data(), c->grp.bperm.size());
  add((void **)&c->d_symconj, c->grp.conj.data(), c->grp.conj.size());
  // goal pre-images of every candidate, and the coefficient limit below which no child leaves int32: a child
  // coefficient is a sum over the gate column of parent coefficients times the column's Z[w] coefficients
  std::vector<int4> xg((size_t)G * N * N, make_int4(0, 0, 0, 0));
  std::vector<uint64_t> xgm(G, 0);
  std::vector<int> xgk(G, -1);
  uint64_t gmax = 1;
  std::vector<HostMat> pre;
  std::vector<char> have;
  c->goal_skip = goal_preimages(s, pre, have);
  for (int g = 0; g < G; g++) {
    const HostGate &h = s.gates[g];
    const HostMat &X = pre[g];
    if (have[g]) {
      xgk[g] = X.k;
      for (int i = 0; i < N * N; i++) {
        const Zq &z = X.e[i];
        xg[(size_t)g * N * N + i] = make_int4((int)z.a, (int)z.b, (int)z.c, (int)z.d);
        if (z.a | z.b | z.c | z.d) xgm[g] |= 1ull << i;
      }
    }
    const int L = 1 << h.nq;
    for (int lc = 0; lc < L; lc++) {
      uint64_t col = 0;
      for (int m = 0; m < L; m++) {
        const Zq &z = h.inv[m * L + lc];
        col += (uint64_t)(std::llabs(z.a) + std::llabs(z.b) + std::llabs(z.c) + std::llabs(z.d));
      }
      gmax = std::max(gmax, col);
    }
  }
  add((void **)&c->d_xg, xg.data(), xg.size() * sizeof(int4));
  add((void **)&c->d_xgm, xgm.data(), xgm.size() * 8);
  add((void **)&c->d_xgk, xgk.data(), xgk.size() * 4);
  c->spec_blob.need(blob.size());
  chz::copy_sync(c->spec_blob.p, blob.data(), blob.size(), hipMemcpyHostToDevice, c->stream);
  for (auto &d : at) *d.first = c->spec_blob.p + d.second;
  KSpec &k = c->ks;
  std::memset(&k, 0, sizeof k);
  k.n = s.n; k.G = G; k.depth = c->depth; k.standard = s.standard; k.lite = s.lite; k.tdiv = s.tdiv;
  k.gk = s.gk; k.nsym = c->grp.size(); k.ncuts = (int)s.cutmask.size(); k.have_table = !s.cxtable.empty();
  for (int i = 0; i < k.ncuts; i++) { k.cutmask[i] = s.cutmask[i]; k.cutsize[i] = s.cutsize[i]; k.rad[i] = s.rad[i]; }
  // a cut bound from lowered needs is a lower bound only if the table never decreases when a need grows
  k.cx_mono = 1;
  k.ncx = (int)s.cxtable.size();
  k.ncd = c->cdv.size() <= (size_t)CD_MAX ? (int)c->cdv.size() : 0;
  for (int j = 0; j < k.ncd; j++) k.cdv[j] = c->cdv[j];
  k.cxlds = s.cxtable.size() <= (size_t)CXL_MAX;
  for (int v : s.cxtable) if (v < 0 || v > 255) k.cxlds = 0;
  if (k.have_table) {
    size_t total = 1;
    for (int i = 0; i < k.ncuts; i++) total *= (size_t)s.rad[i];
    if (total != s.cxtable.size()) k.cx_mono = 0;
    for (size_t idx = 0; k.cx_mono && idx < total; idx++) {
      size_t rest = idx, mult = 1;
      for (int i = 0; i < k.ncuts; i++) {
        const size_t nd = rest % (size_t)s.rad[i];
        rest /= (size_t)s.rad[i];
        if (nd + 1 < (size_t)s.rad[i] && s.cxtable[idx + mult] < s.cxtable[idx]) k.cx_mono = 0;
        mult *= (size_t)s.rad[i];
      }
    }
  }
  // Montgomery images of w^i: w = 3^((p-1)/8) mod p, times 2^32
  auto mulm = [](uint64_t a, uint64_t b) { return (a * b) % P_MOD; };
  uint64_t w = 1, b = 3, e = (P_MOD - 1) / 8;
  while (e) { if (e & 1) w = mulm(w, b); b = mulm(b, b); e >>= 1; }
  const uint64_t R = (1ull << 32) % P_MOD;
  uint64_t wp = 1;
  for (int i = 0; i < 4; i++) { k.wm[i] = (uint32_t)mulm(wp, R); wp = mulm(wp, w); }
  k.gates = c->d_gates; k.bad2 = c->d_bad2; k.bad3 = c->d_bad3; k.goal = c->d_goal; k.cxtable = c->d_cx;
  k.goalnz = 0;
  for (int i = 0; i < N * N; i++) {
    const Zq &z = s.goal[i];
    if (z.a || z.b || z.c || z.d) k.goalnz |= 1ull << i;
  }
  k.symperm = c->d_symperm; k.symconj = c->d_symconj;
  k.xg = c->d_xg; k.xgm = c->d_xgm; k.xgk = c->d_xgk;
  k.maglim = (uint32_t)(0x7fffffffull / gmax);
  k.cjm = 0;
  for (int y = 0; y < k.nsym && y < 16; y++) k.cjm |= (uint32_t)(c->grp.conj[y] != 0) << y;
  std::vector<LevelTabs> tabs(1);
  build_level_tabs(tabs[0], dg, k, c->grp, s);
  c->ktab.need(sizeof(LevelTabs));
  chz::copy_sync(c->ktab.p, tabs.data(), sizeof(LevelTabs), hipMemcpyHostToDevice, c->stream);
  k.ktab = (const uint4 *)c->ktab.p;
}

// a new search of the batch: the depth-0 cases end at once (src/cheese.cu:1794-1801), the rest wait to join
// the frontier as roots at the next step
static int batch_add_parsed(chz_ctx *c, const HostSpec &hs, int depth, int tag) {
  const int N = hs.N(), E = N * N;
  chz_event ev{};
  ev.tag = tag;
  ev.kind = 1;
  for (int i = 0; i < E; i++) {
    const Zq &z = hs.target[i];
    if (z.a != (int32_t)z.a || z.b != (int32_t)z.b || z.c != (int32_t)z.c || z.d != (int32_t)z.d) {
      set_err("target coefficients exceed int32");
      ev.result.rc = 4;
      c->events.push_back(ev);
      return 0;
    }
  }
  HostMat U;
  U.e = hs.target;
  U.k = hs.tk;
  if (!c->nogoal && host_equal(U, c->hs.goal, c->hs.gk)) { ev.result.rc = 0; c->events.push_back(ev); return 0; }
  if (depth <= 0) { ev.result.rc = 1; c->events.push_back(ev); return 0; }
  // a free instance slot: not held by a live search
  std::vector<bool> used(c->nslots + c->binst.size() + 1, false);
  for (const BInst &b : c->binst)
    if (b.live) {
      used[b.slot] = true;
      c->shared_steps = true;      // the batch now steps more than one search
    }
  int slot = 0;
  while (used[slot]) slot++;
  try {
    HIPCHK(hipSetDevice(c->device));
    reserve_slots(c, slot + 1);
  } catch (DeviceError &) {
    return 5;
  }
  int bi = -1;
  for (size_t i = 0; i < c->binst.size(); i++) if (!c->binst[i].live) { bi = (int)i; break; }
  if (bi < 0) { c->binst.emplace_back(); bi = (int)c->binst.size() - 1; }
  BInst &b = c->binst[bi];
  b = BInst{};
  b.tag = tag;
  b.depth = depth;
  b.slot = slot;
  b.live = true;
  b.t0 = now_s();
  Pending q;
  q.inst = bi;
  q.k = hs.tk;
  q.U.resize(E);
  for (int i = 0; i < E; i++) {
    const Zq &z = hs.target[i];
    q.U[i] = make_int4((int)z.a, (int)z.b, (int)z.c, (int)z.d);
  }
  c->pending.push_back(std::move(q));
  return 0;
}

#include "table_join.h"

}  // namespace chz

// ------------------------------------------------------------------------------------------------------
// C ABI

extern "C" {

const char *chz_last_error(void) { return chz::get_err(); }

int chz_open(const char *spec_text, size_t len, const chz_opts *opts, chz_ctx **out) {
  *out = nullptr;
  std::unique_ptr<chz_ctx> c(new chz_ctx());
  std::string err;
  if (const int prc = chz::parse_spec(std::string(spec_text, len), c->hs, err)) { chz::set_err(err); return prc; }
  chz_opts o{-1, 1, 1, 0, 0, 0};
  if (opts) o = *opts;
  c->opts = o;
  c->depth = o.depth >= 0 ? o.depth : c->hs.depth;
  c->ev.mask = o.profile ? ~0u : 0u;
  if (c->depth > CHZ_DMAX) { chz::set_err("depth > 48"); return 3; }
  if (c->hs.states) { chz::set_err("state targets are not supported by this build"); return 3; }
  if (c->hs.n < 1 || c->hs.n > 3) { chz::set_err("n = " + std::to_string(c->hs.n) + " not built (1..3)"); return 3; }
  if (c->hs.truncated) { chz::set_err("malformed spec"); return 2; }
  if ((int)c->hs.target.size() != c->hs.N() * c->hs.N()) { chz::set_err("spec has no target"); return 2; }
  // symmetry classes: not with --mitm, whose join compares exact residuals (src/cheese.cu:1786)
  c->grp = chz::build_symmetry(c->hs, o.sym != 0 && !o.mitm);
  if (o.table < 0 || o.table > CHZ_DMAX) { chz::set_err("--table K outside 0..48"); return 3; }
  // default occupancy: 5 waves per SIMD fits the n = 3 kernel's registers; smaller groups spill there, so 4
  if (c->hs.n < 3)
    for (auto &sc : c->sched) sc.var = 0;
  else
    // n = 3 defaults per level-size bucket (log2 of the parents in a launch), from the autotuner's winners on the
    // Toffoli budget-12 and budget-13 proofs on MI355X: up to 32K parents run at 4 waves per SIMD (the variant with
    // the most registers), below 8K parents with a parent's candidates split over lane groups down to about 2,048
    // groups per launch, 8K-32K in two lane groups per parent; mid-size levels want fewer blocks, the largest more
    for (int b = 0; b < (int)(sizeof c->sched / sizeof c->sched[0]); b++) {
      chz_ctx::Sched &sc = c->sched[b];
      if (b == 13 || b == 14) { sc.var = 0; sc.cap = 4096; sc.tw = 16384; }
      else if (b <= 12) { sc.var = 0; sc.tw = 2048; }
      else if (b >= 15 && b <= 17) sc.cap = 4096;
      else if (b >= 19) sc.cap = 16384;
    }
  try {
    c->device = o.device;
    HIPCHK(hipSetDevice(o.device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    chz::StreamScope scope(c->stream);
    chz::upload_spec(c.get());
  } catch (chz::DeviceError &) {
    return 5;
  }
  *out = c.release();
  return 0;
}

int chz_run(chz_ctx *c, chz_level_cb cb, void *user, chz_result *out) {
  chz_result tmp{};
  chz_result *res = out ? out : &tmp;
  std::memset(res, 0, sizeof *res);
  if (chz_batch_live(c)) { chz::set_err("the context's batch holds searches"); return res->rc = 3; }
  for (int f = 0; f < chz::K_NFAM; f++) { c->k_launch[f] = 0; c->k_ms[f] = 0; c->k_work[f] = 0; }
  for (double &v : c->k_stage) v = 0;
  c->events.clear();
  res->split_kind = 0;
  res->split[0] = res->split[1] = -1;
  if (c->opts.mitm || c->opts.table > 0) {   // --mitm takes precedence, as in the reference (src/cheese.cu:1803)
    try {
      HIPCHK(hipSetDevice(c->device));
      chz::StreamScope scope(c->stream);
      switch (c->hs.n) {
        case 1: return c->opts.mitm ? chz::run_mitm<1>(c, cb, user, res) : chz::run_table<1>(c, cb, user, res);
        case 2: return c->opts.mitm ? chz::run_mitm<2>(c, cb, user, res) : chz::run_table<2>(c, cb, user, res);
        default: return c->opts.mitm ? chz::run_mitm<3>(c, cb, user, res) : chz::run_table<3>(c, cb, user, res);
      }
    } catch (chz::DeviceError &) {
      return res->rc = 5;
    }
  }
  const int rc0 = chz::batch_add_parsed(c, c->hs, c->depth, 0);
  if (rc0) return res->rc = rc0;
  while (true) {
    chz_event e;
    bool done = false;
    while (!c->events.empty()) {
      e = c->events.front();
      c->events.pop_front();
      if (e.kind == 0) { if (cb) cb(&e.level, user); }
      else { *res = e.result; done = true; }
    }
    if (done) return res->rc;
    if (const int rc = chz_batch_step(c)) return res->rc = rc;
  }
}

// the expansion schedule per level-size bucket, chosen by timing every candidate schedule on this context's own
// search (reps runs each); the fastest mean launch time per bucket wins.  Per-level results do not depend on
// the schedule.  out (may be NULL) receives the chosen table as JSON text.
int chz_tune(chz_ctx *c, int reps, char *out, size_t cap) {
  if (chz_batch_live(c)) { chz::set_err("the context's batch holds searches"); return 3; }
  const int vars[3] = {0, 1, 2}, caps[3] = {4096, 8192, 16384}, tws[5] = {2048, 4096, 8192, 16384, 32768};
  std::vector<chz_ctx::Sched> cands;
  for (int v : vars) for (int cp : caps) for (int tw : tws) { chz_ctx::Sched s; s.var = v; s.cap = cp; s.tw = tw; cands.push_back(s); }
  std::vector<std::vector<double>> sum(cands.size(), std::vector<double>(32, 0.0));
  std::vector<std::vector<int>> num(cands.size(), std::vector<int>(32, 0));
  try {
    HIPCHK(hipSetDevice(c->device));
    for (size_t k = 0; k < cands.size(); k++) {
      c->tuning = true;
      c->force = cands[k];
      for (int r = 0; r < std::max(1, reps); r++) {
        chz_result res{};
        const int rc = chz_run(c, nullptr, nullptr, &res);
        if (rc >= 2) { c->tuning = false; return rc; }
        HIPCHK(hipStreamSynchronize(c->stream));
        for (auto &te : c->tune_ev) {
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, te.second.first, te.second.second));
          sum[k][te.first] += ms;
          num[k][te.first]++;
          c->ev.pool.push_back(te.second.first);
          c->ev.pool.push_back(te.second.second);
        }
        c->tune_ev.clear();
      }
    }
  } catch (chz::DeviceError &) {
    c->tuning = false;
    return 5;
  }
  c->tuning = false;
  std::string js = "{";
  bool first = true;
  for (int b = 0; b < 32; b++) {
    int best = -1;
    for (size_t k = 0; k < cands.size(); k++)
      if (num[k][b] && (best < 0 || sum[k][b] / num[k][b] < sum[best][b] / num[best][b])) best = (int)k;
    if (best < 0) continue;
    c->sched[b] = cands[best];
    char buf[128];
    std::snprintf(buf, sizeof buf, "%s\"%d\": {\"waves_per_eu\": %d, \"grid_cap\": %d, \"split_waves\": %d, \"ms\": %.4f}",
                  first ? "" : ", ", b, cands[best].var + 4, cands[best].cap, cands[best].tw, sum[best][b] / num[best][b]);
    js += buf;
    first = false;
  }
  js += "}";
  if (out && cap) { std::strncpy(out, js.c_str(), cap - 1); out[cap - 1] = 0; }
  return 0;
}

int chz_batch_add(chz_ctx *c, const char *spec_text, size_t len, int depth, int tag) {
  chz::HostSpec hs;
  std::string err;
  if (const int prc = chz::parse_spec(std::string(spec_text, len), hs, err)) { chz::set_err(err); return prc; }
  if (hs.truncated) { chz::set_err("malformed spec"); return 2; }
  if (hs.states) { chz::set_err("state targets are not supported by this build"); return 3; }
  if ((int)hs.target.size() != hs.N() * hs.N()) { chz::set_err("spec has no target"); return 2; }
  if (!chz::same_gate_set(hs, c->hs)) { chz::set_err("the spec's gate set differs from the batch's"); return 3; }
  const int d = depth >= 0 ? depth : hs.depth;
  if (d > CHZ_DMAX) { chz::set_err("depth > 48"); return 3; }
  return chz::batch_add_parsed(c, hs, d, tag);
}

int chz_batch_step(chz_ctx *c) {
  try {
    HIPCHK(hipSetDevice(c->device));
    chz::StreamScope scope(c->stream);
    switch (c->hs.n) {
      case 1: chz::batch_step<1>(c); break;
      case 2: chz::batch_step<2>(c); break;
      case 3: chz::batch_step<3>(c); break;
    }
    return 0;
  } catch (chz::DeviceError &) {
    return 5;
  }
}

void chz_batch_set_budget(chz_ctx *c, double bytes) { c->mem_budget = bytes > 0 ? bytes : 0; }

long long chz_batch_waits(const chz_ctx *c) { return c->carried_levels; }

long long chz_batch_frontier(chz_ctx *c, long long max, long long stride, int *paths, int *lens, uint32_t *bw,
                             int *tags) {
  if (stride < 1) stride = 1;
  long long n = 0;
  try {
    HIPCHK(hipSetDevice(c->device));
    chz::StreamScope scope(c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<int> seq;
    for (long long j = 0; j < c->m && n < max; j += stride, n++) {
      uint32_t w = 0, slot = 0;
      chz::copy_sync(&w, c->pb[c->cur].p + j, 4, hipMemcpyDeviceToHost, c->stream);
      chz::copy_sync(&slot, c->finst[c->cur].p + j, 4, hipMemcpyDeviceToHost, c->stream);
      // the node's gates from its root: the tree walked up from the current level, as walk_circuit does
      seq.clear();
      uint32_t i = (uint32_t)j;
      for (int s = c->gstep; s >= 0 && c->tree_par[s]; s--) {
        uint32_t par;
        uint8_t g;
        chz::copy_sync(&par, c->tree_par[s] + i, 4, hipMemcpyDeviceToHost, c->stream);
        if (par == chz::NO_PARENT) break;
        chz::copy_sync(&g, c->tree_gate[s] + i, 1, hipMemcpyDeviceToHost, c->stream);
        if (g != chz::CARRY_GATE) seq.push_back(g);
        i = par;
      }
      std::reverse(seq.begin(), seq.end());
      const int len = (int)std::min<size_t>(seq.size(), CHZ_DMAX);
      for (int t = 0; t < len; t++) paths[(size_t)n * CHZ_DMAX + t] = seq[t];
      lens[n] = len;
      bw[n] = w;
      tags[n] = -1;
      for (const chz::BInst &b : c->binst)
        if (b.live && (uint32_t)b.slot == slot) tags[n] = b.tag;
    }
  } catch (chz::DeviceError &) {
    return -1;
  }
  return n;
}

int chz_batch_events(chz_ctx *c, chz_event *out, int max) {
  int k = 0;
  while (k < max && !c->events.empty()) { out[k++] = c->events.front(); c->events.pop_front(); }
  return k;
}

int chz_batch_live(const chz_ctx *c) {
  int n = 0;
  for (const chz::BInst &b : c->binst) n += b.live;
  return n;
}

long long chz_batch_nodes(const chz_ctx *c) {
  long long n = c->m;
  return n + (long long)c->pending.size();
}

int chz_num_gates(const chz_ctx *c) { return c->hs.G(); }
int chz_num_qubits(const chz_ctx *c) { return c->hs.n; }
int chz_depth(const chz_ctx *c) { return c->depth; }
int chz_sym_size(const chz_ctx *c) { return c->grp.size(); }
const char *chz_gate_name(const chz_ctx *c, int g) {
  return g >= 0 && g < c->hs.G() ? c->hs.gates[g].name.c_str() : "";
}

int chz_profile(const chz_ctx *c, const char **names, double *seconds, int n) {
  int j = 0;
  for (int f = 0; f < chz::K_NFAM && j < n; f++, j++) {
    if (names) names[j] = chz::k_names[f];
    if (seconds) seconds[j] = c->k_ms[f] / 1000.0;
  }
  return j;
}

unsigned chz_time_kernels(chz_ctx *c, unsigned families) {
  const unsigned prev = c->ev.mask;
  c->ev.mask = families;
  return prev;
}

int chz_expand_stages(const chz_ctx *c, double *out) {
  for (int i = 0; i < 6; i++) out[i] = c->k_stage[i];
  return 6;
}

int chz_kernel_stats(const chz_ctx *c, chz_kstat *out, int n) {
  int j = 0;
  for (int f = 0; f < chz::K_NFAM && j < n; f++, j++) {
    out[j].name = chz::k_names[f];
    out[j].launches = c->k_launch[f];
    out[j].ms = c->k_ms[f];
    out[j].work = c->k_work[f];
  }
  return j;
}

static int verify_hs(const chz::HostSpec &hs, const int *circuit, int len) {
  // product of the forward gate matrices, gate 1 first, compared exactly with the target (verify.py:20-33)
  const int n = hs.n, N = 1 << n;
  chz::HostMat P;
  P.e.assign(N * N, chz::Zq{0, 0, 0, 0});
  for (int i = 0; i < N; i++) P.e[i * N + i] = chz::Zq{1, 0, 0, 0};
  for (int j = 0; j < len; j++) {
    if (circuit[j] < 0 || circuit[j] >= hs.G()) return -1;
    // P <- g P via (P^dag g^dag)^dag: apply g on the left with its forward matrix
    const chz::HostGate &g = hs.gates[circuit[j]];
    const int L = 1 << g.nq;
    int mask = 0;
    for (int q = 0; q < g.nq; q++) mask |= 1 << g.q[q];
    std::vector<chz::Zq> out(N * N, chz::Zq{0, 0, 0, 0});
    for (int r = 0; r < N; r++) {
      int lr = 0;
      for (int q = 0; q < g.nq; q++) lr |= ((r >> g.q[q]) & 1) << q;
      const int base = r & ~mask;
      for (int m = 0; m < L; m++) {
        const chz::Zq gm = g.fwd[lr * L + m];
        if (!(gm.a | gm.b | gm.c | gm.d)) continue;
        int x = base;
        for (int q = 0; q < g.nq; q++) if ((m >> q) & 1) x |= 1 << g.q[q];
        for (int col = 0; col < N; col++) {
          const chz::Zq v = P.e[x * N + col];
          chz::Zq &o = out[r * N + col];
          o.a += gm.a * v.a - gm.b * v.d - gm.c * v.c - gm.d * v.b;
          o.b += gm.a * v.b + gm.b * v.a - gm.c * v.d - gm.d * v.c;
          o.c += gm.a * v.c + gm.b * v.b + gm.c * v.a - gm.d * v.d;
          o.d += gm.a * v.d + gm.b * v.c + gm.c * v.b + gm.d * v.a;
        }
      }
    }
    P.e.swap(out);
    P.k += g.gk;
    while (P.k > 0) {
      bool all = true;
      for (auto &z : P.e) if (((z.a - z.c) & 1) || ((z.b - z.d) & 1)) { all = false; break; }
      if (!all) break;
      for (auto &z : P.e) z = chz::Zq{(z.b - z.d) / 2, (z.a + z.c) / 2, (z.b + z.d) / 2, (z.c - z.a) / 2};
      P.k--;
    }
  }
  return chz::host_equal(P, hs.target, hs.tk) ? 1 : 0;
}

int chz_verify(const chz_ctx *c, const int *circuit, int len) { return verify_hs(c->hs, circuit, len); }

int chz_verify_spec(const char *spec_text, size_t len, const int *circuit, int n) {
  chz::HostSpec hs;
  std::string err;
  if (chz::parse_spec(std::string(spec_text, len), hs, err) || hs.truncated ||
      (int)hs.target.size() != hs.N() * hs.N()) {
    chz::set_err(err.empty() ? "malformed spec" : err);
    return -2;
  }
  return verify_hs(hs, circuit, n);
}

void chz_close(chz_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  chz::free_tree(c);
  for (int b = 0; b < 2; b++) {
    c->res[b].release(); c->pk[b].release(); c->g1[b].release(); c->g2[b].release();
    c->pb[b].release();
    c->sk[b].release(); c->sv[b].release();
  }
  c->finst[0].release(); c->finst[1].release(); c->ctl.release(); c->d_active.release(); c->gflag.release();
  c->scan_tmp.release(); c->d_lo.release(); c->d_hi.release();
  c->pin.release();
  for (auto &g : c->sgraph) if (g.exec) { (void)hipGraphExecDestroy(g.exec); g.exec = nullptr; }
  c->status.release(); c->kept.release(); c->keys.release(); c->surv_o.release(); c->surv_key.release();
  c->surv_bw.release();
  c->info.release(); c->bcount.release(); c->hist.release(); c->runs.release(); c->kidx.release();
  c->spec_blob.release();
  c->ktab.release();
  if (c->gt) { c->gt->release(); delete c->gt; c->gt = nullptr; }
  c->gt_keys.release(); c->gt_hit.release(); c->gt_cost.release(); c->gt_best.release();
  c->ev.destroy();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
