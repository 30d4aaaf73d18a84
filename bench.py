"""Benchmark of the exact gate-count search on one MI355X.

A step is one complete search of the headline workload: the Toffoli target (tests/golden/qasm/toffoli.qasm)
at gate budget 12, an exhaustive proof that no circuit of <= 12 gates exists (exit 1), whose per-level log is
checked against the reference CPU build's golden log on every step.  value = prefixes/s (sum of the
per-level "prefixes" / wall time of a step); ms_per_step is the time to proof.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-runs R] [--no-cpu] [--serve N] [--no-table]
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "prefixes/s & time-to-proof, Toffoli d=12; served solves/s + p50/p99 vs plain"
# 256 CUs x 4 SIMD16 x 2.4 GHz: peak 32-bit integer VALU lane operations per second
VALU32_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12


def _strip(lines):
    out = []
    for ln in lines:
        d = json.loads(ln)
        d.pop("seconds", None)
        out.append(d)
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(runs=3, threads=16):
    """The CPU baseline timed on this host beside the GPU measurement: oracle/_ref/cheese-cpu, the reference's own
    src/cheese.cu compiled here as its CPU configuration (Thrust's OpenMP backend; rocThrust's headers stand in for
    CCCL's), or without it the oracle's C++ restatement ("port").  `threads` OpenMP threads (16: the host's CPU
    share of one GPU; nproc is reported beside it), pinned (OMP_PROC_BIND=close, OMP_PLACES=cores); the median of
    `runs` full proofs each of Toffoli budget 9 and budget 12, every log checked against the golden log.  value:
    prefixes/s of the budget-12 proof."""
    ref = os.path.join(ROOT, "oracle", "_ref", "cheese-cpu")
    port = os.path.join(ROOT, "oracle", "cheese-oracle")
    kind, exe = ("reference", ref) if os.path.exists(ref) else ("port", port)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    out = {}
    for d in (9, 12):
        spec = os.path.join(ROOT, "tests", "golden", "specs", f"toffoli-d{d}.spec")
        with open(os.path.join(ROOT, "tests", "golden", "logs", f"toffoli-d{d}.log")) as f:
            golden = _strip(f.read().strip().splitlines())
        times, pref = [], 0
        for _ in range(runs):
            t = time.perf_counter()
            p = subprocess.run([exe, spec], capture_output=True, text=True, env=env)
            times.append(time.perf_counter() - t)
            lines = [l for l in p.stdout.splitlines() if l.strip()]
            if p.returncode != 1 or _strip(lines) != golden:
                print(f"bench.py: CPU baseline log of Toffoli d{d} differs from the golden log", file=sys.stderr)
                sys.exit(1)
            pref = sum(json.loads(l).get("prefixes", 0) for l in lines)
        med = sorted(times)[len(times) // 2]
        out[d] = {"ms_per_proof": round(med * 1e3, 1), "runs_ms": [round(x * 1e3, 1) for x in times],
                  "prefixes": pref}
    d12 = out[12]
    return {"value": round(d12["prefixes"] / (d12["ms_per_proof"] / 1e3), 1), "unit": "prefixes/s",
            "cores": threads, "kind": kind,
            "build": ("reference src/cheese.cu, CPU configuration (Thrust OpenMP backend, rocThrust headers), "
                      "compiled here") if kind == "reference" else "oracle/cheese_oracle.cpp restatement",
            "sample": f"Toffoli proofs at budgets 9 and 12, {runs} runs each, median; value from budget 12",
            "toffoli_d9": out[9], "toffoli_d12": out[12], "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "pinning": "OMP_PROC_BIND=close OMP_PLACES=cores"}


def copy_probe(nbytes):
    """one device-to-device copy of nbytes through the HIP runtime: a kernel of known traffic (nbytes read,
    nbytes written) against which the profiler's FETCH_SIZE / WRITE_SIZE counters are scaled"""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def toffoli_proof(depth, runs=5):
    """a Toffoli impossibility proof at one budget on one context (budget 9: C1; budget 14: the optimum-15 proof),
    log checked against the reference's golden log every run; median wall time per proof"""
    import cheese_amd
    with open(os.path.join(ROOT, "tests", "golden", "specs", f"toffoli-d{depth}.spec")) as f:
        spec = f.read()
    with open(os.path.join(ROOT, "tests", "golden", "logs", f"toffoli-d{depth}.log")) as f:
        golden = _strip(f.read().strip().splitlines())
    s = cheese_amd.Searcher(spec)
    try:
        s.run()
        times, prefixes = [], 0
        for _ in range(runs):
            t = time.perf_counter()
            out = s.run()
            times.append(time.perf_counter() - t)
            if out["rc"] != 1 or _strip([cheese_amd.level_line(l) for l in out["levels"]]) != golden:
                print(f"bench.py: Toffoli d{depth} log differs from the reference golden log", file=sys.stderr)
                sys.exit(1)
            prefixes = sum(l["prefixes"] for l in out["levels"])
    finally:
        s.close()
    ms = sorted(times)[len(times) // 2] * 1e3
    return {"ms_per_proof": round(ms, 3), "prefixes": prefixes, "prefixes_per_s": round(prefixes / (ms / 1e3), 1)}


def gate_table(ks=(7, 8), runs=3):
    """--table K on the Toffoli target (src/cheese.cu:1431-1593, :1804-1842): per K one context whose first run
    builds the distance table (every symmetry class of <= K gates) and proves budget 14 impossible, then the proof
    again with the table kept (median of runs); with the largest K the budget-15 search finds the optimal 15-gate
    circuit, verified exactly.  The reference quotes K=7 2.65M classes, K=8 17.9M, and with --table 8 and its table
    cached budget 14 in 1.5 s and budget 15 in 11.1 s (docs/PLAN.md:87-92)."""
    import cheese_amd
    with open(os.path.join(ROOT, "tests", "golden", "specs", "toffoli-d14.spec")) as f:
        spec14 = f.read()
    spec15 = spec14.replace("\ndepth 14\n", "\ndepth 15\n")
    assert spec15 != spec14
    out = {}
    for K in ks:
        s = cheese_amd.Searcher(spec14, table=K)
        try:
            t = time.perf_counter()
            first = s.run()
            first_s = time.perf_counter() - t
            tlv = [l for l in first["levels"] if l["dir"] == 1]
            times = []
            for _ in range(runs):
                t = time.perf_counter()
                r = s.run()
                times.append(time.perf_counter() - t)
                if r["rc"] != 1 or first["rc"] != 1:
                    print(f"bench.py: Toffoli budget 14 with --table {K} did not prove impossibility", file=sys.stderr)
                    sys.exit(1)
        finally:
            s.close()
        out[f"k{K}"] = {"classes": 1 + sum(l["distinct"] for l in tlv), "first_run_s": round(first_s, 3),
                        "budget14_proof_ms": round(sorted(times)[len(times) // 2] * 1e3, 2)}
    K = max(ks)
    s = cheese_amd.Searcher(spec15, table=K)
    try:
        s.run()
        t = time.perf_counter()
        r = s.run()
        ms = (time.perf_counter() - t) * 1e3
    finally:
        s.close()
    if r["rc"] != 0 or not r["verified"] or len(r["circuit"]) != 15:
        print(f"bench.py: Toffoli budget 15 with --table {K}: no verified 15-gate circuit", file=sys.stderr)
        sys.exit(1)
    out[f"k{K}"]["budget15_found_ms"] = round(ms, 2)
    out[f"k{K}"]["budget15_circuit"] = r["circuit"]
    out["reference_quoted_s"] = {"k7_build": 3.6, "k8_build": 30, "budget14_k8": 1.5, "budget15_k8": 11.1}
    return out


def random_circuits(n=32):
    """the second half of the workload: seeded random 3-qubit Clifford+T circuits of 10..13 gates, each
    synthesised at budget = its length; every found circuit is checked exactly against its target"""
    import cheese_amd
    from cheese_amd import workloads
    total_t, total_p, lens = 0.0, 0, []
    for seed in range(n):
        circ, spec = workloads.random_spec(seed)
        t = time.perf_counter()
        out = cheese_amd.search(spec)
        total_t += time.perf_counter() - t
        if out["rc"] != 0 or not out["verified"] or len(out["circuit"]) > len(circ):
            print(f"bench.py: random circuit {seed} not synthesised exactly", file=sys.stderr)
            sys.exit(1)
        total_p += sum(l["prefixes"] for l in out["levels"])
        lens.append([len(circ), len(out["circuit"])])
    return {"seconds": round(total_t, 3), "prefixes": total_p, "prefixes_per_s": round(total_p / total_t, 1),
            "lengths_input_found": lens}


def clifford_table(T, wall, nq=100000):
    """C4: the whole 3-qubit Clifford group tabulated on the GPU (BFS into the HBM hash table), then 1e5 seeded
    random words of 5..40 gates answered in one batched launch"""
    from cheese_amd import clifford
    entries = T.digest()[1]
    words = clifford.random_words(7, nq, T.ng)
    dist, ms = T.query_words(words)
    return {"elements": entries, "levels": T.levels(), "build_ms_device": round(T.build_ms, 2),
            "build_s_wall": round(wall, 3), "queries": nq, "query_ms_device": round(ms, 3),
            "queries_per_s": round(nq / (ms / 1e3), 1), "max_distance": max(dist)}


def tcount_batch(cliff, count=256):
    """C3: the 3-qubit coset table (K = 4) built once, then `count` seeded targets (k T/Tdg gates, k in 3..7,
    between uniformly random Cliffords from the Clifford table) searched in one batched call; every
    decomposition is verified exactly and its T-count is at most k"""
    from cheese_amd import tcount as T
    from cheese_amd import workloads as W
    with open(os.path.join(ROOT, "tests", "golden", "tspec", "toffoli.tspec")) as f:
        base = f.read()
    t = time.perf_counter()
    table = T.CosetTable(base, K=4)
    build = time.perf_counter() - t
    cases = T.c3_targets(cliff, count)
    t = time.perf_counter()
    specs = [T.tspec_text(3, W.unitary(3, circ)) for _, circ in cases]
    gen = time.perf_counter() - t
    ks0 = table.kernel_stats()
    t = time.perf_counter()
    res = table.search(specs, maxt=16)
    search = time.perf_counter() - t
    ks1 = table.kernel_stats()
    hist = {}
    for (k, circ), r in zip(cases, res):
        if r["rc"] != 0 or r["tcount"] > k or not T.verify(3, W.unitary(3, circ), [list(x) for x in r["rotations"]]):
            print("bench.py: a T-count target was not decomposed exactly", file=sys.stderr)
            sys.exit(1)
        hist[r["tcount"]] = hist.get(r["tcount"], 0) + 1
    out = {"table_cosets": table.size, "table_build_s": round(build, 3),
           "table_levels": [[lv["t"], lv["new"], round(lv["seconds"], 3)] for lv in table.levels],
           "targets": count, "search_s": round(search, 3), "solves_per_s": round(count / search, 1),
           "tcount_histogram": {str(k): v for k, v in sorted(hist.items())}, "verified": count,
           "target_generation_s_host": round(gen, 2),
           "search_kernels_ms": {k: round(ks1[k]["ms"] - ks0[k]["ms"], 3) for k in ks1}}
    table.close()
    return out


def serving(cliff, n_req=400, seed=1):
    """C5: a seeded mixed stream (50% gate-count 2/3-qubit circuits of 4..13 gates at budget = length, 35% T-count
    with T-count <= k in 2..7, 15% Clifford re-synthesis of 5..40-letter words).  Capacity: the whole stream
    offered at once, plain (one request at a time) and served (continuous batching).  Latency: Poisson arrivals
    at 1.5x the plain capacity, the same trace for both modes.  Served answers must equal plain answers."""
    from cheese_amd import serve
    from cheese_amd import workloads as W
    t = time.perf_counter()
    reqs = serve.mixed_requests(n_req, seed, cliff)
    gen = time.perf_counter() - t
    spec2 = W.spec_from_skeleton(W.skeleton(2), W.unitary(2, []), 1)
    spec3 = W.spec_from_skeleton(W.skeleton(3), W.unitary(3, []), 1)
    with open(os.path.join(ROOT, "tests", "golden", "tspec", "toffoli.tspec")) as f:
        tspec = f.read()
    out = {"requests": n_req, "mix": {"gates": sum(r["kind"] == serve.GATES for r in reqs),
                                      "tcount": sum(r["kind"] == serve.TCOUNT for r in reqs),
                                      "clifford": sum(r["kind"] == serve.CLIFFORD for r in reqs)},
           "generation_s_host": round(gen, 2)}
    res = {}
    for mode in ("plain", "served"):
        srv = serve.Server(spec2, spec3, tspec, plain=(mode == "plain"))
        try:
            ans, wall = serve.replay(srv, reqs, [0.0] * n_req)
            if srv.error():
                print(f"bench.py: {mode} server device error: {srv.error()}", file=sys.stderr)
            if mode == "served":
                out["served_engines"] = srv.stats()
                out["served_peak_device_gb"] = round(srv.peak_device_bytes / 2**30, 2)
                out["served_peak_frontier_estimate_gb"] = round(srv.peak_bytes / 2**30, 2)
                out["hbm_gb"] = 288
        finally:
            srv.close()
        res[mode] = ans
        out[mode + "_capacity"] = serve.summarise(ans, wall)
        out[mode + "_finish_by_kind_s"] = serve.finish_by_kind(reqs, ans)
        if mode == "plain":
            out["plain_service_by_kind"] = serve.service_breakdown(reqs, ans)
    bad = []
    for i in range(n_req):
        a, b = res["plain"][i], res["served"][i]
        if any(a[f] != b[f] for f in ("rc", "len", "out", "found", "split")):
            bad.append(i)
    if bad:
        for i in bad[:8]:
            r = reqs[i]
            print(f"bench.py: served answer {i} differs from the plain answer: kind {r['kind']}, n {r.get('n')}, "
                  f"length {r.get('length')}; plain {[res['plain'][i][f] for f in ('rc', 'len', 'out', 'found')]}, "
                  f"served {[res['served'][i][f] for f in ('rc', 'len', 'out', 'found')]}", file=sys.stderr)
        print(f"bench.py: {len(bad)} of {n_req} served answers differ", file=sys.stderr)
        sys.exit(1)
    rate = 1.5 * out["plain_capacity"]["solves_per_s"]
    arrivals = serve.poisson_arrivals(n_req, rate, seed)
    for mode in ("plain", "served"):
        srv = serve.Server(spec2, spec3, tspec, plain=(mode == "plain"))
        try:
            ans, wall = serve.replay(srv, reqs, arrivals)
        finally:
            srv.close()
        out[mode + "_poisson"] = serve.summarise(ans, wall)
    out["poisson_rate_per_s"] = round(rate, 2)
    out["capacity_ratio"] = round(out["served_capacity"]["solves_per_s"] / out["plain_capacity"]["solves_per_s"], 2)
    out["answers_equal"] = True
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-runs", type=int, default=3, help="CPU baseline: runs per budget (median)")
    ap.add_argument("--cpu-threads", type=int, default=16, help="CPU baseline: OpenMP threads (the host's share)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0, help="also run a D2D copy of this size (profiling)")
    ap.add_argument("--random", action=argparse.BooleanOptionalAction, default=True,
                    help="also time the 32 seeded random 3-qubit circuits (budget = length)")
    ap.add_argument("--clifford", action=argparse.BooleanOptionalAction, default=True,
                    help="also build the 3-qubit Clifford table and answer 1e5 queries")
    ap.add_argument("--tcount", action=argparse.BooleanOptionalAction, default=True,
                    help="also build the 3-qubit T-count table and search 256 batched targets")
    ap.add_argument("--d9", action=argparse.BooleanOptionalAction, default=True,
                    help="also time the Toffoli budget-9 and budget-14 proofs")
    ap.add_argument("--table", action=argparse.BooleanOptionalAction, default=True,
                    help="also time --table 7 and --table 8 on the Toffoli budget-14 proof and the budget-15 circuit")
    ap.add_argument("--tune", action=argparse.BooleanOptionalAction, default=True,
                    help="autotune the expansion schedule on the workload before the timed proofs")
    ap.add_argument("--serve", type=int, default=2000, metavar="N",
                    help="also serve a mixed stream of N requests, plain and batched (0: skip)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU sharding is a design note only", file=sys.stderr)
        sys.exit(2)
    import cheese_amd
    if a.copy_probe_mb:
        copy_probe(a.copy_probe_mb << 20)
    name = "toffoli-d12"
    with open(os.path.join(ROOT, "tests", "golden", "specs", name + ".spec")) as f:
        spec = f.read()
    with open(os.path.join(ROOT, "tests", "golden", "logs", name + ".log")) as f:
        golden = _strip(f.read().strip().splitlines())
    # one context for every proof (spec parsed and uploaded once, device buffers kept), as a long-running caller
    # holds it; the timed region is the searches themselves
    searcher = cheese_amd.Searcher(spec)
    tuned = searcher.tune(2) if a.tune else None
    # HIP events around the k_level launches only (the roofline's kernel duration); the other families count
    # launches without events, which cost host time on every level
    searcher.time_kernels(("expand",))
    for _ in range(a.warmup):
        searcher.run()
    times, prefixes, kstats, stages = [], 0, {}, {}
    for _ in range(a.steps):
        t = time.perf_counter()
        out = searcher.run(stats=True)
        times.append(time.perf_counter() - t)
        got = _strip([cheese_amd.level_line(l) for l in out["levels"]])
        if out["rc"] != 1 or got != golden:
            print("bench.py: per-level log differs from the reference golden log", file=sys.stderr)
            sys.exit(1)
        prefixes = sum(l["prefixes"] for l in out["levels"])
        for k, v in out["expand_stages"].items():
            stages[k] = stages.get(k, 0) + v
        for k, v in out["kernel_stats"].items():
            s = kstats.setdefault(k, {"launches": 0, "ms": 0.0, "work": 0.0})
            for f in s:
                s[f] += v[f]
    total = sum(times)
    tb = kstats.get("expand", {"ms": 0.0, "work": 0.0, "launches": 0})
    traffic = valu_busy = None   # from the committed rocprofv3 summaries (FETCH/WRITE_SIZE; SQ_INSTS_VALU)
    tpath = os.path.join(ROOT, "profiles", "k_level-toffoli-d12.traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f)
        traffic = round(prof["k_level_hbm_bytes_per_launch"])
        valu_busy = prof.get("k_level_valu_busy")
    achieved = tb["work"] / (tb["ms"] / 1e3) / 1e12 if tb["ms"] else 0.0
    line = {
        "metric": METRIC,
        "value": round(prefixes * a.steps / total, 1),
        "unit": "prefixes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * total / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "reference target tests/qasm/toffoli.qasm, exact; log checked against the reference each step",
        "config": {"workload": "Toffoli gate-count proof, budget 12 (3 qubits, 27 candidates, 12 symmetries)",
                   "prefixes_per_step": prefixes, "parallelism": "1 GPU"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": VALU32_PEAK_TOPS, "unit": "Top/s",
                     "frac": round(achieved / VALU32_PEAK_TOPS, 4), "traffic": traffic,
                     "kernel": "k_level", "kernel_ms_per_step": round(tb["ms"] / a.steps, 3),
                     "launches_per_step": tb["launches"] / a.steps,
                     "work": "algorithmic 32-bit integer lane ops: stage counts x per-stage ops (residual, "
                             "structural bounds, cut ranks, T bound: its f = 0 part per evaluation and each "
                             "flip-pair pass it ran at two 16-bit ops per lane op, class keys)",
                     "stages_per_step": {k: v // a.steps for k, v in stages.items()},
                     "valu_busy_profiled": valu_busy},
        "kernel_launches_per_step": {k: v["launches"] / a.steps for k, v in kstats.items() if v["launches"]},
        "autotune": {b: [v["waves_per_eu"], v["grid_cap"], v["split_waves"]] for b, v in tuned.items()} if tuned else None,
    }
    if a.d9:
        line["toffoli_d9"] = toffoli_proof(9)
        if os.path.exists(os.path.join(ROOT, "tests", "golden", "logs", "toffoli-d14.log")):
            line["toffoli_d14"] = toffoli_proof(14, runs=3)
    if a.table:
        line["gate_table"] = gate_table()
    if a.random:
        line["random32"] = random_circuits()
    if a.clifford or a.tcount or a.serve:
        from cheese_amd import clifford
        t = time.perf_counter()
        cliff = clifford.CliffordTable(3)
        wall = time.perf_counter() - t
        if a.clifford:
            line["clifford"] = clifford_table(cliff, wall)
        if a.tcount:
            line["tcount256"] = tcount_batch(cliff)
        if a.serve:
            line["serving"] = serving(cliff, a.serve)
        cliff.close()
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a.cpu_runs, a.cpu_threads)
        cpu_ms = line["cpu_baseline"]["toffoli_d12"]["ms_per_proof"]
        line["time_to_proof_vs_cpu"] = {"toffoli_d12": round(cpu_ms / line["ms_per_step"], 1)}
        if "toffoli_d9" in line:
            line["time_to_proof_vs_cpu"]["toffoli_d9"] = round(
                line["cpu_baseline"]["toffoli_d9"]["ms_per_proof"] / line["toffoli_d9"]["ms_per_proof"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
