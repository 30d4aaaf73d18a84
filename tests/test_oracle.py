"""The oracle (CPU restatement of the reference search) pinned against the reference's own outputs and
against what the mathematics fixes.  CPU only."""
import ctypes
import random

import numpy as np
import pytest

from conftest import golden_cases, manifest, oracle_run, read_log, read_spec, strip_seconds

from cheese_amd import workloads


@pytest.mark.parametrize("name", golden_cases(max_seconds=4.0))
def test_oracle_matches_reference_logs(oracle_lib, name):
    """bit-identical per-level lines and result line vs the reference CPU build (tests/golden/logs)"""
    m = manifest()[name]
    if "--table" in m["flags"] or "--mitm" in m["flags"]:
        pytest.skip("the oracle restates the plain level search; --table / --mitm logs are checked on the GPU path "
                    "and by the published table sizes")
    kw = dict(depth=-1, dedupe="--no-dedupe" not in m["flags"], sym="--no-sym" not in m["flags"])
    rc, out = oracle_run(oracle_lib, read_spec(name), **kw)
    assert rc == m["rc"]
    assert strip_seconds(out) == strip_seconds(read_log(name))


@pytest.mark.parametrize("name", ["toffoli-d11-nosym", "rus13-unitary-d13"])
def test_oracle_matches_reference_logs_large(oracle_lib, name):
    """the same on larger fixtures (5-40 s each here): the largest --no-sym proof and the lite bound (a ccx gate
    set); the wide T transform's fixture (t-ladder-d42, target sqrt2 exponent 14) is among the fast ones above"""
    m = manifest()[name]
    kw = dict(depth=-1, dedupe="--no-dedupe" not in m["flags"], sym="--no-sym" not in m["flags"])
    rc, out = oracle_run(oracle_lib, read_spec(name), **kw)
    assert rc == m["rc"]
    assert strip_seconds(out) == strip_seconds(read_log(name))


def test_published_ccz_d8_frontier():
    """docs/ALGORITHM.md:191: CCZ 8 largest frontier 4,960 exact residuals vs 501 symmetry classes; the golden
    logs (with and without --no-sym) agree"""
    assert max(l["distinct"] for l in strip_seconds(read_log("ccz-d8"))) == 501
    assert max(l["distinct"] for l in strip_seconds(read_log("ccz-d8-nosym"))) == 4960


def test_published_toffoli_d11_nosym_frontier():
    """docs/ALGORITHM.md:191: Toffoli 11 largest frontier 172,054 exact residuals (--no-sym)"""
    assert max(l["distinct"] for l in strip_seconds(read_log("toffoli-d11-nosym"))) == 172054


def test_lite_fixtures_use_the_lite_bound():
    """the rus-13 unitary's specs carry `lite 1` (a ccx gate set), so their golden logs pin the lite bound"""
    for d in (12, 13, 14):
        assert "\nstandard 0\nlite 1\ntdiv 2\n" in read_spec(f"rus13-unitary-d{d}")


@pytest.mark.parametrize("name,classes", [("toffoli-d9-table4", 5429), ("toffoli-d10-table5", 45924)])
def test_published_table_sizes(name, classes):
    """docs/ALGORITHM.md:220: 3-qubit Clifford+T distance tables hold 5,429 classes within 4 gates and 45,924
    within 5 (the identity and each table level's new classes); the golden table logs agree"""
    lines = [l for l in strip_seconds(read_log(name)) if l.get("dir") == "table"]
    assert 1 + sum(l["distinct"] for l in lines) == classes


@pytest.mark.parametrize("mode", ["table", "mitm"])
def test_table_and_join_verdicts_equal_the_plain_search(mode):
    """--table and --mitm answer what the plain search answers: the same exit code and circuit length"""
    m = manifest()
    for name, v in m.items():
        if f"--{mode}" not in v["flags"] or v.get("rc") is None:
            continue
        base = name.rsplit("-", 1)[0]
        if base not in m or m[base].get("rc") is None:
            continue
        assert v["rc"] == m[base]["rc"], name
        if v["rc"] == 0:
            a = [l for l in strip_seconds(read_log(name)) if "circuit" in l][0]
            b = [l for l in strip_seconds(read_log(base)) if "circuit" in l][0]
            assert len(a["circuit"]) == len(b["circuit"]), name


def test_published_toffoli_d11_frontier():
    """docs/ALGORITHM.md:191: Toffoli 11 largest frontier 49,503 classes; the golden log agrees"""
    lines = strip_seconds(read_log("toffoli-d11"))
    assert max(l["distinct"] for l in lines) == 49503


def test_published_toffoli_d11_nosym_level6():
    """docs/PLAN.md:43-44: Toffoli 11 --no-sym, level 6: 221,049 survivors -> 147,656 distinct"""
    lines = strip_seconds(read_log("toffoli-d11-nosym"))
    l6 = [l for l in lines if l["t"] == 6][0]
    assert (l6["survivors"], l6["distinct"]) == (221049, 147656)


def test_reference_verdicts():
    """tests/check.sh:16-22 and tests/bench.sh:7 verdicts"""
    m = manifest()
    want = {"cs-d4": 1, "cs-d5": 0, "toffoli-d8": 1, "ccz-d7": 1, "qft_2-d7": 0, "qft_2-d6": 1,
            "relative-phase-toffoli-d9": 0, "toffoli-d9": 1, "toffoli-d10": 1, "ccz-d8": 1,
            "half_adder_3-d9": 1, "fredkin-d9": 1, "peres-d9": 1, "grover_3-d10": 1}
    for k, v in want.items():
        assert m[k]["rc"] == v, k


# ---------- single-matrix bounds ----------

def _mat(entries):
    flat = [c for z in entries for c in z]
    return (ctypes.c_longlong * len(flat))(*flat)


def _residual(L, spec, seq):
    n = int(spec.split("\nn ")[1].split()[0])
    E = 1 << (2 * n)
    buf = (ctypes.c_longlong * (4 * E))()
    arr = (ctypes.c_int * max(1, len(seq)))(*seq)
    k = L.orc_residual(spec.encode(), len(spec), arr, len(seq), buf)
    assert k >= 0
    return n, k, buf


def _complex(buf, n):
    w = np.exp(1j * np.pi / 4)
    E = 1 << (2 * n)
    v = np.array([buf[4 * i] + buf[4 * i + 1] * w + buf[4 * i + 2] * w ** 2 + buf[4 * i + 3] * w ** 3 for i in range(E)])
    return v.reshape(1 << n, 1 << n)


@pytest.mark.parametrize("specname,n", [("cs-d4", 2), ("toffoli-d9", 3)])
def test_t_bound_wht_equals_definition(oracle_lib, specname, n):
    """the WHT form of the T bound equals the definitional PTM sde (the reference's CHECK_T cross-check,
    src/cheese.cu:712-714) on residuals of random prefixes"""
    spec = read_spec(specname)
    G = int(spec.split("\ngates ")[1].split()[0])
    rng = random.Random(7)
    for _ in range(40 if n == 3 else 150):
        seq = [rng.randrange(G) for _ in range(rng.randrange(0, 9))]
        n_, k, buf = _residual(oracle_lib, spec, seq)
        assert oracle_lib.orc_bound(n_, k, buf, 0, 0) == oracle_lib.orc_bound(n_, k, buf, 1, 0)


def _single(n, gates):
    U = workloads.unitary(n, gates)
    k, M = U
    return k, [z for row in M for z in row]


def test_closed_form_bounds(oracle_lib):
    """PTM sde of single gates (python/cheese/spec.py ptm_sde; docs/ALGORITHM.md: T 1, ccx 2), support,
    H structure of permutations"""
    L = oracle_lib
    k, e = _single(1, [("t", (0,))])
    assert L.orc_bound(1, k, _mat(e), 0, 0) == 1
    k, e = _single(3, [("ccx", (0, 1, 2))])
    assert L.orc_bound(3, k, _mat(e), 0, 0) == 2
    assert L.orc_bound(3, k, _mat(e), 1, 0) == 2
    assert L.orc_bound(3, k, _mat(e), 2, 0) == 3          # acts on all three qubits
    assert L.orc_bound(3, k, _mat(e), 3, 0) == 2          # non-affine permutation needs 2 H
    k, e = _single(3, [("cx", (0, 2))])
    assert L.orc_bound(3, k, _mat(e), 2, 0) == 2
    assert L.orc_bound(3, k, _mat(e), 3, 0) == 0          # affine
    assert L.orc_bound(3, k, _mat(e), 0, 0) == 0          # Clifford
    k, e = _single(3, [])
    assert L.orc_bound(3, k, _mat(e), 2, 0) == 0
    k, e = _single(2, [("h", (0,)), ("t", (1,)), ("h", (1,)), ("t", (1,))])
    assert L.orc_bound(2, k, _mat(e), 0, 0) == 2          # two T gates that do not merge


@pytest.mark.parametrize("specname", ["toffoli-d9", "cs-d4"])
def test_fp_rank_equals_complex_rank(oracle_lib, specname):
    """operator-Schmidt rank over F_p (p = 998244353) equals the complex rank on random residuals"""
    spec = read_spec(specname)
    G = int(spec.split("\ngates ")[1].split()[0])
    rng = random.Random(3)
    for _ in range(60):
        seq = [rng.randrange(G) for _ in range(rng.randrange(0, 10))]
        n, k, buf = _residual(oracle_lib, spec, seq)
        S = _complex(buf, n)
        N = 1 << n
        for A in range(1, 1 << n):
            sa = bin(A).count("1")
            if sa > n // 2:
                continue
            # operator-Schmidt matrix: rows (o_A, i_A), cols (o_B, i_B)
            rows = {}
            M = np.zeros((4 ** sa, N * N // 4 ** sa), dtype=complex)
            for o in range(N):
                for i in range(N):
                    oa = ia = ob = ib = na = nb = 0
                    for q in range(n):
                        if (A >> q) & 1:
                            oa |= ((o >> q) & 1) << na; ia |= ((i >> q) & 1) << na; na += 1
                        else:
                            ob |= ((o >> q) & 1) << nb; ib |= ((i >> q) & 1) << nb; nb += 1
                    M[(oa << na) | ia, (ob << nb) | ib] = S[o, i]
            exact = np.linalg.matrix_rank(M, tol=1e-6)
            assert oracle_lib.orc_bound(n, k, buf, 4, A) == exact


# ---------- spec writing and workloads ----------

@pytest.mark.parametrize("name", sorted(manifest()))
def test_spec_writer_reproduces_reference_spec(name):
    """the skeleton + exact-target writer gives the reference spec writer's text byte for byte"""
    m = manifest()[name]
    if "circuit" in m:
        n, circ = 3, [(g, tuple(q)) for g, q in m["circuit"]]
    else:
        import os
        from conftest import ROOT
        path = os.path.join(ROOT, "tests", "golden", "qasm", m["qasm"])
        with open(path) as f:
            n, circ = workloads.parse_qasm(f.read())
        extra = {g for g, _ in circ} - set(workloads.CLIFFORD_T)
        if extra:
            pytest.skip(f"gate set beyond the default skeleton ({', '.join(sorted(extra))})")
    text = workloads.spec_from_skeleton(workloads.skeleton(n), workloads.unitary(n, circ), m["depth"])
    assert text == read_spec(name)


def test_random_circuits_are_seeded():
    for name, m in manifest().items():
        if "seed" in m:
            circ = workloads.random_circuit(m["seed"], 3, 10, 13)
            assert [[g, list(q)] for g, q in circ] == m["circuit"]
            assert m["depth"] == len(circ)


def test_invariances_the_expansion_carries(oracle_lib):
    """A child keeps its parent's T bound when its gate is Clifford (the Pauli transfer matrix of a Clifford gate
    is a signed permutation), and keeps the F_p rank of every cut its gate does not straddle (a local invertible
    factor leaves an operator-Schmidt rank unchanged).  T and T-dagger children do change the T bound."""
    L = oracle_lib
    spec = read_spec("toffoli-d9")
    names = [l.split()[0] for l in spec.splitlines() if l.startswith("L_")]
    qubits = [[int(q) for q in nm.split("_")[2:]] for nm in names]
    clifford = [nm.split("_")[1] not in ("t", "tdg") for nm in names]
    G = len(names)
    cuts = [1, 2, 4]
    rng = random.Random(11)
    changed = 0
    for _ in range(25):
        seq = [rng.randrange(G) for _ in range(rng.randrange(0, 8))]
        n, k, buf = _residual(L, spec, seq)
        t0 = L.orc_bound(n, k, buf, 0, 0)
        r0 = [L.orc_bound(n, k, buf, 4, A) for A in cuts]
        for g in range(G):
            n1, k1, b1 = _residual(L, spec, seq + [g])
            t1 = L.orc_bound(n1, k1, b1, 0, 0)
            if clifford[g]:
                assert t1 == t0, (seq, names[g])
            else:
                assert abs(t1 - t0) <= 1
                changed += t1 != t0
            mask = sum(1 << q for q in qubits[g])
            for A, r in zip(cuts, r0):
                if (mask & A) == 0 or (mask & ~A) == 0:
                    assert L.orc_bound(n1, k1, b1, 4, A) == r, (seq, names[g], A)
    assert changed > 0


def test_alive_budget_closed_form():
    """the expansion decides "alive at T bound t" by t <= rem - h - max(c2, supp - rem) instead of evaluating the
    bound combination (src/cheese.cu:599-607: the minimum over c in [c2, max(c2, supp)] of c + max(h + t,
    supp - 2c)) against rem; the two agree for every h, t, c2, supp and rem >= 0 in range"""
    def combine(h, t, c2, supp):
        return min(c + max(h + t, supp - 2 * c) for c in range(c2, max(c2, supp) + 1))

    for h in range(12):
        for c2 in range(8):
            for supp in range(8):
                for rem in range(16):
                    budget = rem - h - max(c2, supp - rem)
                    for t in range(24):
                        assert (combine(h, t, c2, supp) <= rem) == (t <= budget), (h, t, c2, supp, rem)
