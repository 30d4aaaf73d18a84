/* cheese_hip.h - C ABI of the MI355X (gfx950) exact branch-and-bound synthesis library libcheese_hip.so.
 *
 * The reference has no in-process API: its boundary is a process contract (a .spec text file in, one JSON
 * line per level on stdout, exit codes).  These entry points are that contract made callable: every return
 * code is the reference binary's exit code for the same situation, and every level record carries the
 * fields of the reference's per-level JSON line.  Plain pointers and sizes only; the caller owns its input
 * buffers, a context owns its device memory.  A context is used by one host thread at a time.
 *
 * Return codes (all entry points): 0 circuit found / ok, 1 proof (no circuit within the budget), 2 bad
 * spec, 3 usage or unsupported input, 4 coefficient overflow or capacity, 5 device error.
 */
#ifndef CHEESE_HIP_H
#define CHEESE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHZ_DMAX 48 /* longest prefix, src/cheese.cu:53 */

/* ---------------- gate-count search (src/cheese.cu) ---------------- */

typedef struct chz_ctx chz_ctx;

typedef struct chz_opts {
  int depth;      /* < 0: the spec's depth line; replaces `--depth D` (src/cheese.cu:1861) */
  int dedupe;     /* 0 = `--no-dedupe` (src/cheese.cu:1862) */
  int sym;        /* 0 = `--no-sym` (src/cheese.cu:1866) */
  long chunk;     /* parents per expansion pass, `--chunk N` (src/cheese.cu:1863); <= 0: default */
  int profile;    /* nonzero: per-pass seconds kept for chz_profile (`--profile`, src/cheese.cu:1230-1247) */
  int device;     /* HIP device ordinal */
  int table;      /* K > 0: `--table K` (src/cheese.cu:1804-1842), the distance table of the last K gates (built on
                     the context's first run, kept with it) */
  int mitm;       /* nonzero: `--mitm` (src/cheese.cu:1664-1760), the meet-in-the-middle join; symmetry classes off */
} chz_opts;

/* one level of the search: the fields of the per-level JSON line (src/cheese.cu:1354-1356) */
typedef struct chz_level {
  int t;
  long parents, prefixes, cut_w3, cut_lb, survivors, distinct;
  int proved_unsat;
  double seconds;
  int dir;        /* 0 the search; 1 "dir": "table" (the table build), 2 "fwd", 3 "bwd" (--mitm sides); 4 the join
                     line {"dir": "join", "forward": parents, "backward": prefixes, "hash_matches": distinct,
                     "seconds"} (src/cheese.cu:1714-1715) */
} chz_level;

typedef void (*chz_level_cb)(const chz_level *lvl, void *user);

/* result of a run (src/cheese.cu:1658-1662, 1800, 1846-1848) */
typedef struct chz_result {
  int rc;                  /* 0 found, 1 proof, 4 overflow, 5 device error */
  int len;                 /* circuit length t (0 with rc 0: the target is the goal at depth 0) */
  long found;              /* children equal to the goal at that level ("found" field) */
  int circuit[CHZ_DMAX];   /* candidate indices, in circuit order */
  double seconds;          /* since the search started */
  int split_kind;          /* 1 --table: the "table": [split[0], split[1]] field of a circuit (prefix levels, table
                              gates), or with rc 1 the proof line {"depth": D, "proved_unsat": true, "table": K}
                              (split[1] = K); 2 --mitm: "mitm": [a, b] of a circuit or of the proof line; 0 none */
  int split[2];
} chz_result;

/* Parses a .spec (read_spec, src/cheese.cu:1149-1226), builds the symmetry group (sym_group, :1379-1429) and
 * uploads the spec (run, :1762-1792).  Returns 0, 2 (bad spec) or 3 (n outside 1..3, states mode, depth > 48). */
int chz_open(const char *spec_text, size_t len, const chz_opts *opts, chz_ctx **out);

/* Runs the level-synchronous search (levels + run, src/cheese.cu:1249-1377, :1794-1848).  cb (may be NULL)
 * is called once per completed level, on the calling thread, before the next level starts. */
int chz_run(chz_ctx *ctx, chz_level_cb cb, void *user, chz_result *out);

int chz_num_gates(const chz_ctx *ctx);
int chz_num_qubits(const chz_ctx *ctx);
int chz_depth(const chz_ctx *ctx);
int chz_sym_size(const chz_ctx *ctx);                /* "symmetry classes: %d group elements" (:1792) */
const char *chz_gate_name(const chz_ctx *ctx, int g); /* candidate names L_<gate>_<q>... */
void chz_close(chz_ctx *ctx);

/* Per-pass device seconds of the last run (`--profile` report, src/cheese.cu:1244-1247): names and values
 * for up to n passes; returns the number of passes. */
int chz_profile(const chz_ctx *ctx, const char **names, double *seconds, int n);

/* Device-side kernel statistics of the last run (no reference counterpart; for measurement): per kernel
 * family, launches, summed HIP-event milliseconds (0 for a family that chz_time_kernels leaves untimed) and the
 * algorithmic work it did. */
typedef struct chz_kstat {
  const char *name;
  long launches;
  double ms;
  double work;    /* algorithmic units: integer ops for the expansion kernels, bytes for sort/dedupe */
} chz_kstat;
int chz_kernel_stats(const chz_ctx *ctx, chz_kstat *out, int n);

/* kernel families, in chz_kernel_stats order */
enum { CHZ_FAM_EXPAND = 0, CHZ_FAM_COMPACT, CHZ_FAM_SORT, CHZ_FAM_DEDUPE, CHZ_FAM_GATHER, CHZ_FAM_LEVEL_GRAPH };

/* Which kernel families record HIP events around their launches (no reference counterpart; for measurement): bit f
 * = family f.  chz_open sets every family when opts->profile is nonzero (`--profile` reads these times through
 * chz_profile) and none otherwise: an event pair per launch costs host time on every level, 0.4 ms of a 9 ms
 * Toffoli budget-12 proof with every family timed.  Launch counts are kept either way.  Returns the previous mask;
 * takes effect at the next launch. */
unsigned chz_time_kernels(chz_ctx *ctx, unsigned families);

/* Stage counts of the last run's expansion (no reference counterpart; for measurement), the factors of its work
 * model: out[0] residuals formed, [1] children through the structural bounds, [2] cut ranks computed, [3] T-bound
 * evaluations, [4] class keys, [5] flip-pair passes of the T-bound evaluations (0 to 2 each: a pass runs only while
 * the child is undecided).  out holds 6 values; returns 6. */
int chz_expand_stages(const chz_ctx *ctx, double *out);

/* Autotuner (no reference counterpart): times every candidate expansion schedule (occupancy variant x grid cap x
 * candidate-split threshold) on this context's own search, reps runs each, and keeps the fastest per level-size
 * bucket (log2 of the parents in a launch); later runs on the context use that table.  Results do not depend on
 * the schedule.  out (may be NULL) receives the table as JSON.  0, 3 (batch busy), 4/5 as chz_run. */
int chz_tune(chz_ctx *ctx, int reps, char *out, size_t cap);

/* Exact check of a gate list against the spec's target (verify.py:20-33 restated): 1 equal, 0 not, <0 error. */
int chz_verify(const chz_ctx *ctx, const int *circuit, int len);
/* the same check against the target of a .spec text, on the host only (no device work): 1, 0, <0 error */
int chz_verify_spec(const char *spec_text, size_t len, const int *circuit, int n);

/* last error message of the calling thread (stderr text of the reference) */
const char *chz_last_error(void);

/* ---- batches (no reference counterpart: the reference runs one search per process) ----
 * A context also carries a batch: any number of searches whose specs share its gate set (everything but the
 * target and the depth, e.g. every spec written for the same qubit count and gate set) advance level by level
 * in the same kernel launches, their frontier nodes tagged by search.  Each search's level lines and result
 * are those it gets alone (src/cheese.cu:1249-1377 per search).  chz_run drives a batch of one. */
typedef struct chz_event {
  int tag;             /* the caller's tag of the search */
  int kind;            /* 0: a level line, 1: the search's result (it has finished) */
  chz_level level;
  chz_result result;
} chz_event;
/* queues a search; it joins at the next step.  depth < 0: the spec's depth.  0, 2 (bad spec), 3 (other gate
 * set, states mode, depth > 48).  A target that is the goal, or depth 0, finishes at once (a result event). */
int chz_batch_add(chz_ctx *ctx, const char *spec_text, size_t len, int depth, int tag);
/* one level of every search in the batch; events are queued.  0, or 5 (device error) */
int chz_batch_step(chz_ctx *ctx);
/* pops up to max queued events in order; returns how many */
int chz_batch_events(chz_ctx *ctx, chz_event *out, int max);
int chz_batch_live(const chz_ctx *ctx);         /* searches not finished */
long long chz_batch_nodes(const chz_ctx *ctx);  /* frontier nodes held on the device, queued roots included */
/* per-level memory planning (no reference counterpart): with bytes > 0, each step expands the longest prefix of
 * the batch's searches (in admission order, at least one) whose planned bytes - children's records and the next
 * frontier by each search's last growth - fit beside the frontier held; the other searches wait, their nodes
 * carried to the next level unexpanded (their level lines and results are unchanged, only later).  0: no limit. */
void chz_batch_set_budget(chz_ctx *ctx, double bytes);
long long chz_batch_waits(const chz_ctx *ctx);  /* search-levels that waited so far */
/* Test access to the frontier between steps (no reference counterpart; for tests): up to max nodes of the
 * batch's current frontier, every stride-th from the first.  Per node: its gates from its search's root
 * (paths[i * CHZ_DMAX + t], lens[i] of them), its bound word (bits 0-7 the exact T bound or 255 when unknown,
 * bits 8 + 3c the need of cut c, bit 31 set when the needs are known) and its search's tag.  Returns the number
 * of nodes written, -1 on a device error. */
long long chz_batch_frontier(chz_ctx *ctx, long long max, long long stride, int *paths, int *lens, uint32_t *bw,
                             int *tags);

/* ---------------- Clifford-group table (src/clifford.cu) ----------------
 * The whole n-qubit Clifford group (n <= 3, up to global phase) tabulated by breadth-first search over the gate
 * set {H, S, Sdg, X, Z on each qubit; CX on each ordered pair} (gate_set, src/clifford.cu:147-155) into an
 * open-addressing hash table in HBM: stabiliser tableau -> minimal gate count.  Letters of a word are indices
 * into that gate set, applied in order. */
typedef struct chz_cliff chz_cliff;

/* BFS table (src/clifford.cu:251-280): 0, 3 (n outside 1..3), 4 (group order mismatch), 5 (device error) */
int chz_cliff_build(int n, int device, chz_cliff **out);
/* number of elements at each gate count ("clifford gates d: count", src/clifford.cu:275); returns levels */
int chz_cliff_levels(const chz_cliff *c, long *counts, int maxd);
int chz_cliff_num_gates(const chz_cliff *c);
double chz_cliff_build_ms(const chz_cliff *c);
/* order-independent checksum of all (tableau, distance) entries and their number (no reference counterpart:
 * compares a whole table with another in O(1) memory) */
int chz_cliff_digest(const chz_cliff *c, uint64_t *digest, long long *entries);
/* batched distances of words letters[offs[i] .. offs[i+1]) (k_find, src/clifford.cu:207-210); dist -1: absent */
int chz_cliff_query_words(chz_cliff *c, const uint8_t *letters, const long long *offs, long long nq, int *dist,
                          double *ms);
uint64_t chz_cliff_word_tableau(const chz_cliff *c, const uint8_t *letters, int len);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  chz_cliff_query_words with DEVICE buffers d_letters,
 * d_offs (nq + 1 offsets) and d_dist, calling no hipMalloc/hipFree: the words' tableaux are staged in a fixed
 * per-device scratch array of 2^20 entries (8 MiB, held until the process ends), so nq is at most 2^20 (else
 * 4).  Kernels are enqueued on stream (a hipStream_t; NULL: the table's own stream, synchronised before the
 * call returns).  Calls that overlap on one GPU share the scratch and race.  0, 3, 4, 5. */
int chz_cliff_query_words_noalloc(chz_cliff *c, const uint8_t *d_letters, const long long *d_offs, long long nq,
                                  int *d_dist, void *stream);
/* optimal words of tableaux by walking distances down, first gate in gate order (src/clifford.cu:354-372);
 * words: nq x maxlen letters, lens[i] = length or -1 */
int chz_cliff_readback(chz_cliff *c, const uint64_t *tab, long long nq, int maxlen, uint8_t *words, int *lens);
/* count uniformly random group elements (uniform over the table's entries), seeded (no reference counterpart:
 * generates the Clifford factors of random T-count targets); 0 or 4 */
int chz_cliff_sample(chz_cliff *c, uint64_t seed, long long count, uint64_t *tab);
/* `cliffopt JOB [--beam B]` (src/clifford.cu:168-206, :285-385): the job text (n, the T gates in time order as
 * "qubit dagger" lines, the t + 1 Clifford segments as "m gate q [q2] ..." lines, gates h s sdg x z y cx) is
 * rewritten with the same T gates, each moved through a Clifford K (K Z_q' K^-1 = +-Z_q) chosen by a beam search of
 * width beam (1..255) so that the summed optimal segment lengths are small, and every segment replaced by an
 * optimal gate list from the table.  out receives the reference's stdout: {"clifford_gates": ...,
 * "clifford_gates_before": ..., "t_gates": ..., "seconds": ...} and the gate lines.  The placements are taken in
 * the table's sorted element order, so equal-cost choices are deterministic.  0, 2 (bad job), 3 (qubit count
 * differs from the table's), 4 (out too small), 5. */
int chz_cliff_opt(chz_cliff *c, const char *job, size_t len, int beam, char *out, size_t cap);
void chz_cliff_free(chz_cliff *c);
const char *chz_cliff_last_error(void);

/* ---------------- optimal T-count (src/tcount.cu) ----------------
 * Input is the .tspec text (channel representation of U and the Pauli rotation tables, src/tcount.cu:433-450).
 * A table holds every right coset reachable from the identity with <= K rotations, each at the level of its
 * T-count; a search peels rotations off U level by level and looks each residual up (src/tcount.cu:452-532).
 * One table serves any number of searches for the same qubit count; a search call runs a batch of targets
 * together, level by level, each with its own result. */
#define CHZ_TMAX 32 /* deepest rotation sequence, src/tcount.cu:44 */

typedef struct chz_tc chz_tc;

/* one level line: {"dir": ..., "t", "parents", "children", "cosets", "new"} (src/tcount.cu:385-390) */
typedef struct chz_tc_level {
  int t;
  long parents, children, cosets, fresh; /* fresh: the "new" field */
  double seconds;
} chz_tc_level;

typedef struct chz_tc_result {
  int rc;                        /* 0 found, 1 none with T-count <= maxt ("tcount_above"), 2 bad spec,
                                    3 unsupported, 4 overflow */
  int tcount;                    /* with rc 0 */
  int split[2];                  /* rotations peeled off U, table level of the residual */
  int rotations[2 * CHZ_TMAX];   /* tcount pairs (p, inv) in circuit order (the "rotations" field) */
  int nlevels;                   /* search level lines before the result line */
  chz_tc_level levels[CHZ_TMAX + 1];
  double seconds;
} chz_tc_result;

/* builds the table for the spec's qubit count with levels <= K from the identity (the table part of main,
 * src/tcount.cu:473-495): 0, 2 (bad spec), 3 (n outside 1..3, K outside 0..28), 4 (overflow), 5 (device) */
int chz_tc_table(const char *spec_text, size_t len, int K, int device, chz_tc **out);
int chz_tc_table_levels(const chz_tc *t, chz_tc_level *levels, int max); /* returns the number of levels */
long long chz_tc_table_size(const chz_tc *t);                           /* cosets in the table, identity included */
int chz_tc_table_depth(const chz_tc *t);                                 /* levels built (D.top()) */
int chz_tc_num_qubits(const chz_tc *t);
/* searches ninst targets (the search part of main, src/tcount.cu:497-531, `--max maxt`); results per target.
 * Returns 0 when the batch ran (see res[i].rc), 3 on bad arguments, 5 on a device error. */
int chz_tc_search(chz_tc *t, const char *const *specs, const size_t *lens, int ninst, int maxt, chz_tc_result *res);
/* device-side kernel statistics of the table build and every search so far (no reference counterpart) */
int chz_tc_kernel_stats(const chz_tc *t, chz_kstat *out, int n);
/* dedupe work so far (no reference counterpart; for measurement): runs of equal key prefixes, children visited
 * in them, exact comparisons, duplicates removed; returns 4 */
int chz_tc_dedupe_counts(const chz_tc *t, double *out);
void chz_tc_free(chz_tc *t);
const char *chz_tc_last_error(void);

/* Host only (no device work).  The .tspec text of the n-qubit unitary entries / sqrt2^k (entries: N x N
 * row-major, four int64 coefficients a b c d of a + b w + c w^2 + d w^3 each): the normalised transfer matrix
 * and the rotation tables (the writer the reference runs as `python3 -m cheese.tcount spec`,
 * python/cheese/tcount.py:79-129, :157-167).  Writes a NUL-terminated text into out (cap bytes) and returns its
 * length; out NULL: only the length.  -2 not unitary, -3 bad arguments, -4 out too small. */
int chz_tspec_write(int n, int k, const long long *entries, char *out, size_t cap);
/* Host only.  Exact check of a T-count answer (`python3 -m cheese.tcount verify`, python/cheese/tcount.py:
 * 170-183): peels the t rotations (p, inv) pairs off U in order and tests that a Clifford remains.
 * 1 verified, 0 not, <0 bad arguments or not unitary. */
int chz_tc_verify(int n, int k, const long long *entries, const int *rotations, int t);


/* ---------------- request server (no reference counterpart) ----------------
 * Gate-count, T-count and Clifford re-synthesis requests submitted from any thread and answered by worker
 * threads.  Served mode batches continuously: each gate-count batch (2 and 3 qubits) advances a level at a
 * time and admits queued searches between levels while its frontier-memory estimate stays in budget; queued
 * T-count targets are searched together; queued Clifford words are read back together; the workers' kernels
 * overlap on the device.  Plain mode answers one request at a time in arrival order on the same engines.
 * Every answer is the one the request gets alone. */
enum { CHZ_REQ_GATES = 0, CHZ_REQ_TCOUNT = 1, CHZ_REQ_CLIFFORD = 2 };

typedef struct chz_server chz_server;

typedef struct chz_server_opts {
  int device;
  int plain;             /* 1: one request at a time, in arrival order */
  int tcount_K;          /* T-count table depth (<= 0: 4) */
  int max_batch;         /* most live searches in one batch (<= 0: 1024 for 2 qubits, 64 for 3) */
  double mem_budget_gb;  /* admission budget of a gate-count batch's frontier (<= 0: 200) */
} chz_server_opts;

typedef struct chz_request {
  int kind;              /* CHZ_REQ_* */
  uint64_t id;
  const char *text;      /* .spec (gate count) or .tspec (T-count); copied at submission */
  size_t len;
  int budget;            /* gate count: depth (< 0: the spec's); T-count: maximum T-count (< 0: 16) */
  const uint8_t *word;   /* Clifford: letters over the 3-qubit gate set, applied in order; copied */
  int wlen;
} chz_request;

typedef struct chz_response {
  uint64_t id;
  int kind, rc;          /* rc as the matching entry point's */
  int len;               /* gates of the circuit, T-count, or letters of the optimal word */
  int out[2 * CHZ_DMAX]; /* candidate indices / (p, inv) pairs / letters */
  long found;            /* gate count: children equal to the goal at the last level */
  int split[2];          /* T-count: rotations peeled, table level */
  double t_submit, t_done; /* seconds on the server's clock */
} chz_response;

/* spec2 / spec3: any gate-count spec of the 2- and 3-qubit gate sets served (NULL: not served); tspec3: any
 * 3-qubit .tspec (its rotation tables; NULL: no T-count).  Builds the T-count and Clifford tables. */
int chz_server_open(const char *spec2, size_t len2, const char *spec3, size_t len3, const char *tspec3, size_t lent,
                    const chz_server_opts *opts, chz_server **out);
int chz_server_submit(chz_server *s, const chz_request *r);
int chz_server_poll(chz_server *s, chz_response *out, int max);   /* finished answers, up to max */
long long chz_server_outstanding(const chz_server *s);            /* submitted and not yet answered */
long long chz_server_peak_bytes(const chz_server *s);             /* largest frontier-memory estimate seen */
long long chz_server_peak_device_bytes(const chz_server *s);      /* largest device memory in use, measured */
/* per-engine work so far (call between bursts): out[0..1] level steps of the 2- and 3-qubit gate batches,
 * [2..3] their seconds, [4..5] their kernel milliseconds in expansion launches and captured level steps,
 * [6..7] their most live searches, [8] T-count batches, [9] their seconds, [10] Clifford batches, [11] their
 * seconds; returns the number written */
int chz_server_stats(chz_server *s, double *out, int n);
const char *chz_server_error(chz_server *s);                       /* last device error a worker met ("" none) */
double chz_server_clock(const chz_server *s);                      /* seconds since the server opened */
void chz_server_close(chz_server *s);                              /* finishes outstanding work first */

#ifdef __cplusplus
}
#endif
#endif
