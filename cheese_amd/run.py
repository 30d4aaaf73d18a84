"""Fewest-gate synthesis over increasing budgets: the reference's outer loop (run.sh:1-64) over this library.

  python -m cheese_amd.run TARGET.qasm                 optimum: budgets 1, 2, 3, ... until a circuit is found
  python -m cheese_amd.run TARGET.qasm --from 7        same, from budget 7 (when depths < 7 are known impossible)
  python -m cheese_amd.run TARGET.qasm --depth 9       one budget: a circuit of <= 9 gates, or a proof
  python -m cheese_amd.run TARGET.spec ...             a prepared .spec (any gate set): its depth line is the
                                                       budget; the loop ends at the spec's own depth
  options: --out FILE.qasm (default results/<name>-opt.qasm), --no-dedupe, --no-sym, --chunk N, --results DIR

Each budget's per-level JSON lines go to <results>/<name>-d<D>.log, as the reference's loop writes them.  A budget
that ends with exit 1 is a proof (no circuit of <= D gates) and the loop goes on; exit 0 stops it: the circuit is
re-checked exactly against the target (chz_verify_spec, exact Z[w] product on the host, as
python/cheese/verify.py:20-45 does) and written as OpenQASM 2.0.  Any other exit ends the loop with that code.

A QASM target is turned into a .spec by cheese_amd/specgen.py, the restated writer (python/cheese/spec.py): the
default candidate set (cx h s sdg t tdg x z) plus the target's own gates.  Targets on more than three qubits end
with exit 3 (the search covers one to three).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

from . import level_line, result_line, search, specgen, verify_spec
from . import workloads as W


def gates_of(names):
    """candidate names L_<gate>_<q>[_<q>...] -> (gate, qubits) (python/cheese/verify.py:12-17)"""
    out = []
    for nm in names:
        _, g, *qs = nm.split("_")
        out.append((g, tuple(int(q) for q in qs)))
    return out


def qasm_text(n, names) -> str:
    """OpenQASM 2.0 of a found circuit, gate 1 first (python/cheese/verify.py:39-43)"""
    body = "".join(f"{g} " + ",".join(f"q[{q}]" for q in qs) + ";\n" for g, qs in gates_of(names))
    return f"OPENQASM 2.0;\ninclude \"qelib1.inc\";\nqreg q[{n}];\n" + body


def with_depth(spec_text: str, depth: int) -> str:
    """the spec with its depth line set to `depth` (the only budget-dependent line of a .spec)"""
    lines = spec_text.splitlines(keepends=True)
    for i, ln in enumerate(lines):
        if ln.startswith("depth "):
            lines[i] = f"depth {depth}\n"
            return "".join(lines)
    raise ValueError("spec has no depth line")


def spec_depth(spec_text: str) -> int:
    for ln in spec_text.splitlines():
        if ln.startswith("depth "):
            return int(ln.split()[1])
    raise ValueError("spec has no depth line")


def spec_qubits(spec_text: str) -> int:
    for ln in spec_text.splitlines():
        if ln.startswith("n "):
            return int(ln.split()[1])
    raise ValueError("spec has no qubit line")


def target_spec(path: str):
    """(name, qubits, gate count of the input or None, base spec text) of a .qasm or .spec target"""
    name = os.path.splitext(os.path.basename(path))[0]
    with open(path) as f:
        text = f.read()
    if path.endswith(".spec"):
        return name, spec_qubits(text), None, text
    n, circ = W.parse_qasm(text)
    if n not in (1, 2, 3):
        raise NotImplementedError(f"{n} qubits: the search covers one to three")
    return name, n, len(circ), specgen.spec_text(n, circ, max(1, len(circ)))


def run(path: str, first: int = 1, only: int | None = None, out: str | None = None, results: str = "results",
        dedupe: bool = True, sym: bool = True, chunk: int = 0, echo=print) -> int:
    """the budget loop; returns the exit code of run.sh (0 found, 1 none within the budgets, else the search's)"""
    try:
        name, n, ngates, base = target_spec(path)
    except NotImplementedError as e:
        echo(f"unsupported target: {e}")
        return 3
    except (OSError, ValueError) as e:
        echo(f"cannot read {path}: {e}")
        return 2
    last = only if only is not None else (ngates if ngates is not None else spec_depth(base))
    d = only if only is not None else first
    out = out or os.path.join(results, f"{name}-opt.qasm")
    os.makedirs(results, exist_ok=True)
    echo(f"{name}: {ngates if ngates is not None else '?'} gates in the input, budgets {d}..{last}")
    t0 = time.time()
    while d <= last:
        spec = with_depth(base, d)
        s = time.time()
        r = search(spec, depth=d, dedupe=dedupe, sym=sym, chunk=chunk)
        dt = time.time() - s
        log = os.path.join(results, f"{name}-d{d}.log")
        with open(log, "w") as f:
            for lv in r["levels"]:
                f.write(level_line(lv) + "\n")
            if r["rc"] == 0:
                f.write(result_line(r) + "\n")
        if r["rc"] == 0:
            names = r["circuit"] or []
            echo(f"budget {d}: circuit with {len(names)} gates found ({dt:.1f} s)")
            ok = verify_spec(spec, r["circuit_idx"])
            echo(f"{'equivalent' if ok else 'NOT EQUIVALENT'}: {len(names)} gates"
                 + (f" vs {ngates} in the input" if ngates is not None else ""))
            if not ok:
                return 5
            with open(out, "w") as f:
                f.write(qasm_text(n, names))
            echo(f"written {out}")
            if only is None:
                echo(f"optimum: {len(names)} gates (every smaller budget proven impossible above), "
                     f"total {time.time() - t0:.1f} s")
            return 0
        if r["rc"] == 1:
            echo(f"budget {d}: no circuit of <= {d} gates (proof, {dt:.1f} s)")
        else:
            echo(f"budget {d}: open or error (exit {r['rc']}), see {log}")
            return r["rc"]
        d += 1
    if only is None:
        echo("no circuit within the budgets tried")
    return 1


def main(argv=None) -> int:
    a = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    a.add_argument("target")
    a.add_argument("--from", dest="first", type=int, default=1)
    a.add_argument("--depth", type=int)
    a.add_argument("--out")
    a.add_argument("--results", default="results")
    a.add_argument("--no-dedupe", action="store_true")
    a.add_argument("--no-sym", action="store_true")
    a.add_argument("--chunk", type=int, default=0)
    o = a.parse_args(argv)
    return run(o.target, o.first, o.depth, o.out, o.results, not o.no_dedupe, not o.no_sym, o.chunk,
               echo=lambda m: print(m, flush=True))


if __name__ == "__main__":
    sys.exit(main())
