"""T-count oracle pinned against the reference's own T-count logs (benchmarks/optimal/*-tcount.json), the
published table size for 2 qubits, and exact verification of decompositions.  CPU only (the 3-qubit table
takes minutes on a CPU; it is checked on the GPU side against the published size)."""
import ctypes
import json
import os

import pytest

from conftest import GOLDEN, ROOT

from cheese_amd import tcount as T
from cheese_amd import workloads as W

TC = os.path.join(GOLDEN, "tcount")
NAMES2 = ["cs", "qft_2"]
NAMES3 = ["relative-phase-toffoli", "toffoli", "ccz", "peres", "fredkin"]


@pytest.fixture(scope="session")
def tco(tc_oracle):
    return tc_oracle


def _strip(text):
    out = []
    for l in text.strip().splitlines():
        d = json.loads(l)
        d.pop("seconds", None)
        out.append(d)
    return out


def test_two_qubit_table_size_and_searches(tco):
    """n=2 table with K=6: 1,728,736 cosets (docs/PLAN.md:99-100); cs and qft_2 logs as the reference's"""
    spec = open(os.path.join(GOLDEN, "tspec", "cs.tspec")).read().encode()
    buf = ctypes.create_string_buffer(1 << 16)
    h = tco.orc_tc_table(spec, len(spec), 6, 0, buf, len(buf))
    assert tco.orc_tc_table_size(h) == 1728736
    assert _strip(buf.value.decode()) == _strip(open(os.path.join(TC, "table-n2-k6.oracle.log")).read())
    for name in NAMES2:
        s = open(os.path.join(GOLDEN, "tspec", name + ".tspec")).read().encode()
        rot = (ctypes.c_int * 64)()
        rc = tco.orc_tc_search(h, s, len(s), 16, buf, len(buf), rot)
        assert rc == 0
        assert _strip(buf.value.decode()) == _strip(open(os.path.join(TC, name + ".json")).read())
    tco.orc_tc_free(h)


@pytest.mark.parametrize("name", NAMES2 + NAMES3)
def test_tspec_writer_matches_reference(name):
    n, circ = W.parse_qasm(open(os.path.join(TC, name + ".qasm")).read())
    assert T.tspec_text(n, W.unitary(n, circ)) == open(os.path.join(GOLDEN, "tspec", name + ".tspec")).read()


@pytest.mark.parametrize("name", NAMES2 + NAMES3)
def test_reference_decompositions_verify(name):
    """the reference's own decompositions pass the exact verifier; without one rotation they cannot (the
    T-count is optimal, so a shorter list never leaves a Clifford)"""
    n, circ = W.parse_qasm(open(os.path.join(TC, name + ".qasm")).read())
    U = W.unitary(n, circ)
    res = [json.loads(l) for l in open(os.path.join(TC, name + ".json")) if '"rotations"' in l][0]
    assert len(res["rotations"]) == res["tcount"]
    assert T.verify(n, U, res["rotations"])
    for drop in range(len(res["rotations"])):
        assert not T.verify(n, U, res["rotations"][:drop] + res["rotations"][drop + 1:])


def test_residue_exactness_bound():
    """the HIP T-count path identifies channel entries (x + y sqrt2) / sqrt2^e by residues mod p = 2^31 - 1,
    where 2^16 is a square root of 2 and 2^15 its inverse.  For an orthogonal matrix (and hence its orthogonal
    Galois conjugate) |x| <= sqrt2^e and |y| <= sqrt2^(e-1), so x + 2^16 y lies strictly inside (-p/2, p/2) for
    every e <= 28, where the residue fixes the integer, and |x| < 2^15 then fixes (x, y).  At e = 29 the
    bound fails, which is why the host refuses such runs."""
    import math
    p = 2 ** 31 - 1
    assert pow(2 ** 16, 2, p) == 2 and (2 ** 15 * 2 ** 16) % p == 1

    def span(e):
        return math.floor(2 ** (e / 2)) + 2 ** 16 * math.floor(2 ** ((e - 1) / 2))
    assert all(span(e) < p / 2 for e in range(1, 29))
    assert span(29) > p / 2
    seen = {}
    for x in range(-300, 301, 7):
        for y in range(-300, 301, 11):
            r = (x + 2 ** 16 * y) % p
            assert seen.setdefault(r, (x, y)) == (x, y)


@pytest.mark.parametrize("name", NAMES2 + NAMES3)
def test_reference_channels_satisfy_the_entry_bound(name):
    """every entry of the reference's channel matrices obeys |x + y sqrt2| <= sqrt2^e and |x - y sqrt2| <=
    sqrt2^e, and the matrix is exactly orthogonal"""
    toks = open(os.path.join(GOLDEN, "tspec", name + ".tspec")).read().split()
    n = int(toks[toks.index("n") + 1])
    P = 4 ** n
    i = toks.index("channel")
    e = int(toks[i + 1])
    vals = list(map(int, toks[i + 2:i + 2 + 2 * P * P]))
    X, Y = vals[0::2], vals[1::2]
    s2 = 2 ** 0.5
    for x, y in zip(X, Y):
        assert abs(x + y * s2) <= s2 ** e + 1e-9 and abs(x - y * s2) <= s2 ** e + 1e-9
    for b in range(P):
        for c in range(P):
            A = sum(X[a * P + b] * X[a * P + c] + 2 * Y[a * P + b] * Y[a * P + c] for a in range(P))
            B = sum(X[a * P + b] * Y[a * P + c] + Y[a * P + b] * X[a * P + c] for a in range(P))
            assert B == 0 and A == (2 ** e if b == c else 0)


def test_clifford_check_closed_forms():
    """the decomposition check on known answers: a Clifford word needs no rotation; T on qubit q is the rotation
    about Z_q (p = 3 * 4^q), peeled either way round (T = R(Z) I = R(Z)^-1 S); T alone is no Clifford"""
    from cheese_amd import tcount as T
    cliff = [("h", (0,)), ("cx", (0, 2)), ("s", (1,)), ("h", (2,)), ("cx", (2, 1)), ("sdg", (0,)), ("x", (1,))]
    assert T.verify(3, W.unitary(3, cliff), [])
    for q in range(3):
        U = W.unitary(3, [("t", (q,))])
        assert not T.verify(3, U, [])
        assert T.verify(3, U, [(3 << (2 * q), 0)]) and T.verify(3, U, [(3 << (2 * q), 1)])
        assert not T.verify(3, U, [(1 << (2 * q), 0)])          # the rotation about X_q leaves a non-Clifford
    U = W.unitary(3, cliff + [("t", (2,))] + cliff)
    assert not T.verify(3, U, [])
