"""The .spec writer restated (cheese_amd/specgen.py, python/cheese/spec.py:127-226, :356-415): for every golden
spec whose target the reference wrote from a QASM file or a seeded circuit, the text is byte-identical, the
non-default gate set of ricotta-rus-0003 (cz, y) included.  CPU only."""
import ctypes
import os

import pytest

from conftest import GOLDEN, manifest, read_spec
from test_host_spec import _local, _spec, shim  # noqa: F401  (shim: module fixture)

from cheese_amd import run as R
from cheese_amd import specgen as S
from cheese_amd import workloads as W


def _cases():
    out = []
    for name, m in sorted(manifest().items()):
        if m.get("qasm") and not m.get("flags"):
            out.append(name)
    return out


@pytest.mark.parametrize("name", _cases())
def test_spec_of_a_qasm_target(name):
    m = manifest()[name]
    with open(os.path.join(GOLDEN, "qasm", m["qasm"])) as f:
        n, circ = W.parse_qasm(f.read())
    assert S.spec_text(n, circ, m["depth"]) == read_spec(name)


@pytest.mark.parametrize("seed", [0, 7, 21])
def test_spec_of_a_seeded_circuit(seed):
    circ, spec = W.random_spec(seed)
    assert S.spec_text(3, circ, len(circ)) == read_spec(f"random3-s{seed}") == spec


def test_cut_tables():
    """closed forms: one cut (radix 3) for two qubits, table[need] = need; three single-qubit cuts for three"""
    cuts, rad, table = S.cx_table(2)
    assert cuts == [(0,)] and rad == [3] and table == [0, 1, 2]
    cuts, rad, table = S.cx_table(3)
    assert cuts == [(0,), (1,), (2,)] and rad == [3, 3, 3] and len(table) == 27
    assert table[0] == 0 and table[1 + 3 + 9] == 2 and table[2 + 2 * 3 + 2 * 9] == 3


def test_window_rules_cut_cancelling_pairs():
    """a gate followed by its inverse equals the empty product: every such pair is a W2 cut"""
    cs = S.candidates(2, sorted(W.CLIFFORD_T))
    bad2, _ = S.window_rules(cs, use3=False)
    bad = set(bad2)
    for i, (_, g, qs) in enumerate(cs):
        j = next(j for j, (_, h, q2) in enumerate(cs) if h == S.DAGGER.get(g, g) and q2 == qs)
        assert (i, j) in bad


@pytest.mark.parametrize("g", sorted(S.PTM_SDE))
def test_ptm_exponents_match_the_host(shim, g):  # noqa: F811
    nq, gk, M = _local(g)
    spec = _spec(nq, gk, M).encode()
    out = (ctypes.c_int * 8)()
    assert shim.shim_gate_info(spec, len(spec), 0, out, 8) == nq
    assert out[1] == S.PTM_SDE[g]


def test_lite_spec_for_a_ccx_target():
    """a target with ccx: the bounds of the lite mode, T bound divided by the largest PTM exponent (2)"""
    text = S.spec_text(3, [("ccx", (0, 1, 2)), ("h", (2,))], 3)
    assert "standard 0\nlite 1\ntdiv 2\n" in text
    assert "L_ccx_0_1_2 3 0 1 2 0" in text
    assert text.endswith("end\n")


def test_loop_writes_specs_for_other_gates(tmp_path):
    """the budget loop takes targets with gates outside the default set through the writer (no GPU needed to
    build the spec)"""
    q = tmp_path / "r.qasm"
    q.write_text('OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[2];\ny q[0];\ncz q[0],q[1];\n')
    name, n, ng, base = R.target_spec(str(q))
    assert (name, n, ng) == ("r", 2, 2)
    assert base == S.spec_text(2, [("y", (0,)), ("cz", (0, 1))], 2)


@pytest.mark.parametrize("n", [2, 3, 4])
def test_cut_table_by_brute_force(n):
    """the cut table against a direct minimum over bounded per-placement gate counts: for every need vector,
    the fewest gates among all count vectors (each count 0..4) whose crossings reach every need; every need
    vector for 2 and 3 qubits, 400 seeded ones of the 10,125 for 4"""
    import itertools
    import random
    cuts, rad, table = S.cx_table(n)
    pairs = list(itertools.combinations(range(n), 2))
    best = {}
    for counts in itertools.product(range(5), repeat=len(pairs)):
        cross = tuple(sum(c for (a, b), c in zip(pairs, counts) if (a in A) != (b in A)) for A in cuts)
        tot = sum(counts)
        if best.get(cross, 99) > tot:
            best[cross] = tot
    vs = range(len(table)) if n < 4 else random.Random(4).sample(range(len(table)), 400)
    for v in vs:
        need, x = [], v
        for r in rad:
            need.append(x % r)
            x //= r
        want = min(t for cr, t in best.items() if all(c >= q for c, q in zip(cr, need)))
        assert table[v] == want, (n, need)
