"""Writes the golden fixtures of the T-count mode.  Runs only where the reference tree is mounted: it imports
the reference's stdlib-only .tspec writer (python/cheese/tcount.py) for the benchmark targets and copies the
reference's own T-count logs (benchmarks/optimal/*-tcount.json) and target circuits next to them.

Outputs (committed data): tests/golden/tspec/<name>.tspec, tests/golden/tcount/<name>.json, <name>.qasm."""
import io
import os
import shutil
import sys

REF = "/root/reference"
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(REF, "python"))
from cheese import spec as ref_spec  # noqa: E402  (reference, this container only)
from cheese import tcount as ref_tcount  # noqa: E402

HERE = os.path.join(ROOT, "tests", "golden")
OPT = os.path.join(REF, "benchmarks", "optimal")
NAMES = ["cs", "qft_2", "relative-phase-toffoli", "toffoli", "ccz", "peres", "fredkin"]


def main():
    for name in NAMES:
        q = os.path.join(OPT, f"{name}-tcount.qasm")
        n, circ = ref_spec.parse_qasm(q)
        buf = io.StringIO()
        ref_tcount.write_spec(buf, n, ref_spec.unitary(n, circ))
        with open(os.path.join(HERE, "tspec", name + ".tspec"), "w") as f:
            f.write(buf.getvalue())
        shutil.copy(os.path.join(OPT, f"{name}-tcount.json"), os.path.join(HERE, "tcount", name + ".json"))
        shutil.copy(q, os.path.join(HERE, "tcount", name + ".qasm"))
        print(name, n, len(buf.getvalue()))


if __name__ == "__main__":
    main()
