"""Writes the golden fixtures of the gate-count search.  Runs only where the reference tree is mounted
(/root/reference): it writes the .spec files with cheese_amd/specgen.py (the restated writer, byte-identical to
the reference's, as tests/test_specgen.py checks) and runs the reference CPU build (oracle/_ref/cheese-cpu,
built by `make -C oracle ref`) for their per-level logs.

Gate sets with ccx (the rus-13 unitary, the lite bound's fixtures) had no golden spec to pin the restated writer
on, so their .spec files are written by the reference's own stdlib-only writer (python/cheese/spec.py
write_spec), imported here only; tests/test_specgen.py then checks the restated writer against them.

Outputs (committed data): tests/golden/specs/*.spec, tests/golden/logs/*.log, tests/golden/manifest.json,
and cheese_amd/skeletons/n{2,3}.spec (the default gate set's spec with an identity target, from which
targets are spliced in).  Usage: python tests/golden/make_fixtures.py [--only NAME ...]
"""
import json
import os
import subprocess
import sys
import time

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from cheese_amd import specgen, workloads  # noqa: E402

HERE = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "oracle", "_ref", "cheese-cpu")
QASM = os.path.join(HERE, "qasm")      # the reference's tests/qasm targets, committed as data

# (name, qasm file, depth, extra flags): the reference's own regression and timing cases
# (tests/check.sh:16-22, tests/bench.sh:7) plus deeper Toffoli budgets and flag variants
CASES = [
    ("cs-d4", "cs.qasm", 4, []), ("cs-d5", "cs.qasm", 5, []),
    ("qft_2-d6", "qft_2.qasm", 6, []), ("qft_2-d7", "qft_2.qasm", 7, []),
    ("toffoli-d8", "toffoli.qasm", 8, []), ("toffoli-d9", "toffoli.qasm", 9, []),
    ("toffoli-d10", "toffoli.qasm", 10, []), ("toffoli-d11", "toffoli.qasm", 11, []),
    ("toffoli-d12", "toffoli.qasm", 12, []),
    ("toffoli-d13", "toffoli.qasm", 13, []), ("toffoli-d14", "toffoli.qasm", 14, []),
    ("toffoli-d9-nosym", "toffoli.qasm", 9, ["--no-sym"]),
    ("toffoli-d11-nosym", "toffoli.qasm", 11, ["--no-sym"]),
    ("toffoli-d8-nodedupe", "toffoli.qasm", 8, ["--no-dedupe"]),
    ("toffoli-d10-chunk1000", "toffoli.qasm", 10, ["--chunk", "1000"]),
    ("ccz-d7", "ccz.qasm", 7, []), ("ccz-d8", "ccz.qasm", 8, []),
    ("relative-phase-toffoli-d9", "relative-phase-toffoli.qasm", 9, []),
    ("fredkin-d9", "fredkin.qasm", 9, []), ("peres-d9", "peres.qasm", 9, []),
    ("half_adder_3-d9", "half_adder_3.qasm", 9, []), ("grover_3-d10", "grover_3.qasm", 10, []),
    ("ricotta-rus-0003-d10", "ricotta-rus-0003.qasm", 10, []),
    ("ricotta-rus-0003-d11", "ricotta-rus-0003.qasm", 11, []),
    ("ccz-d8-nosym", "ccz.qasm", 8, ["--no-sym"]),
    # benchmarks/rus/13/unitary-opt.qasm: two ccx, so the spec has `lite 1` and the search the lite bound
    ("rus13-unitary-d12", "rus13-unitary-opt.qasm", 12, []),
    ("rus13-unitary-d13", "rus13-unitary-opt.qasm", 13, []),
    ("rus13-unitary-d14", "rus13-unitary-opt.qasm", 14, []),
    # the distance table of the last K gates (--table K): table build lines, the search's levels 1..D-K, then a
    # circuit from the table or the table's proof line
    ("toffoli-d9-table4", "toffoli.qasm", 9, ["--table", "4"]),
    ("toffoli-d10-table5", "toffoli.qasm", 10, ["--table", "5"]),
    ("toffoli-d12-table6", "toffoli.qasm", 12, ["--table", "6"]),
    ("cs-d5-table3", "cs.qasm", 5, ["--table", "3"]),
    ("qft_2-d7-table3", "qft_2.qasm", 7, ["--table", "3"]),
    ("relative-phase-toffoli-d9-table4", "relative-phase-toffoli.qasm", 9, ["--table", "4"]),
    ("peres-d9-table5", "peres.qasm", 9, ["--table", "5"]),
    # the meet-in-the-middle join (--mitm): forward and backward levels, the join line, a circuit or the proof line
    ("toffoli-d9-mitm", "toffoli.qasm", 9, ["--mitm"]),
    ("toffoli-d10-mitm", "toffoli.qasm", 10, ["--mitm"]),
    ("cs-d5-mitm", "cs.qasm", 5, ["--mitm"]),
    ("qft_2-d7-mitm", "qft_2.qasm", 7, ["--mitm"]),
    ("relative-phase-toffoli-d9-mitm", "relative-phase-toffoli.qasm", 9, ["--mitm"]),
    ("ccz-d7-mitm", "ccz.qasm", 7, ["--mitm"]),
    ("fredkin-d9-mitm", "fredkin.qasm", 9, ["--mitm"]),
]
# (H T) on each of the 3 qubits, 8 times (48 gates): target sqrt2 exponent 14, so the T bound of the first
# levels' children runs with n + k > 16 (the wide transform)
T_LADDER = [(g, (q,)) for _ in range(8) for q in range(3) for g in ("h", "t")]
LADDER_CASES = [("t-ladder-d42", 42), ("t-ladder-d44", 44)]
NRANDOM = 32   # seeded random 3-qubit Clifford+T circuits of 10..13 gates, budget = length


def spec_text(n, circ, depth):
    if any(g == "ccx" for g, _ in circ):
        return reference_writer_text(n, circ, depth)
    return specgen.spec_text(n, circ, depth)


def reference_writer_text(n, circ, depth):
    """the reference's .spec writer (this container only)"""
    import io
    sys.path.insert(0, os.path.join(REF, "python"))
    from cheese import spec as ref_spec
    buf = io.StringIO()
    ref_spec.write_spec(buf, n, circ, depth)
    return buf.getvalue()


REF_TIMEOUT = 150   # seconds of reference CPU time per case; slower cases keep their spec without a log
LONG = {"toffoli-d13": 1200, "toffoli-d14": 2400, "t-ladder-d44": 300, "rus13-unitary-d14": 600}    # the Toffoli optimum-15 proofs: run to the end


def run_ref(path, flags, timeout=REF_TIMEOUT):
    import tempfile
    t = time.time()
    try:
        # a fresh working directory: no table cached by an earlier run (build/table-*.bin), so the table is built
        # and its level lines printed
        with tempfile.TemporaryDirectory() as cwd:
            p = subprocess.run([BIN, path] + flags, capture_output=True, text=True, timeout=timeout, cwd=cwd)
    except subprocess.TimeoutExpired:
        return None, None, time.time() - t
    return p.returncode, p.stdout, time.time() - t


def main():
    only = set(sys.argv[sys.argv.index("--only") + 1:]) if "--only" in sys.argv else None
    man_path = os.path.join(HERE, "manifest.json")
    manifest = json.load(open(man_path)) if os.path.exists(man_path) else {}
    for n in (2, 3):
        with open(os.path.join(workloads.SKELETON_DIR, f"n{n}.spec"), "w") as f:
            f.write(spec_text(n, [], 0))
    jobs = []
    for name, q, d, flags in CASES:
        with open(os.path.join(QASM, q)) as f:
            n, circ = workloads.parse_qasm(f.read())
        jobs.append((name, n, circ, d, flags, {"qasm": q}))
    for name, d in LADDER_CASES:
        jobs.append((name, 3, T_LADDER, d, [], {"circuit": [[g, list(q)] for g, q in T_LADDER]}))
    for s in range(NRANDOM):
        circ = workloads.random_circuit(s, 3, 10, 13)
        jobs.append((f"random3-s{s}", 3, circ, len(circ), [], {"seed": s, "circuit": [[g, list(q)] for g, q in circ]}))
    force = "--force" in sys.argv
    for name, n, circ, d, flags, extra in jobs:
        if only and name not in only:
            continue
        text = spec_text(n, circ, d)
        sp = os.path.join(HERE, "specs", name + ".spec")
        with open(sp, "w") as f:
            f.write(text)
        lp = os.path.join(HERE, "logs", name + ".log")
        if name in manifest and os.path.exists(lp) and not force:
            continue
        rc, out, sec = run_ref(sp, flags, LONG.get(name, REF_TIMEOUT))
        if rc is not None:
            with open(lp, "w") as f:
                f.write(out)
        manifest[name] = dict(extra, n=n, depth=d, flags=flags, rc=rc, ref_seconds=round(sec, 2))
        print(f"{name}: rc {rc}, {sec:.1f} s", flush=True)
        with open(man_path, "w") as f:
            json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
