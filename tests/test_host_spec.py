"""Host analysis of candidate gates (cheese_amd/csrc/spec.cpp, built here with g++ through tests/host_shim.cpp):
the Pauli-transfer-matrix exponent that bounds how far a child's T bound can move from its parent's, the Clifford
test (exponent 0) under which the T bound is inherited exactly, and the per-cut operator-Schmidt ranks that bound
how far a cut need can drop.  The oracle is the checker: its T bound and F_p cut rank of the gate's own matrix.
CPU only."""
import ctypes
import os
import subprocess

import pytest

from conftest import read_spec

from cheese_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# local matrices as {(row, col): w-power} with a sqrt2 exponent (bit j of an index = local qubit j)
EXTRA = {
    "cs": (0, {(0, 0): 0, (1, 1): 0, (2, 2): 0, (3, 3): 2}),
    "ct": (0, {(0, 0): 0, (1, 1): 0, (2, 2): 0, (3, 3): 1}),
    "ch": (1, {(1, 1): 0, (1, 3): 0, (3, 1): 0, (3, 3): 4}),     # control-0 block set to sqrt2 I in _local
}


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = tmp_path_factory.mktemp("shim") / "libhostshim.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-o", str(out),
                    os.path.join(HERE, "host_shim.cpp"), os.path.join(ROOT, "cheese_amd", "csrc", "spec.cpp")],
                   check=True)
    L = ctypes.CDLL(str(out))
    L.shim_goal_preimage.restype = ctypes.c_int
    L.shim_goal_preimage.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                                     ctypes.c_int]
    L.shim_goal_skip.restype = ctypes.c_int
    L.shim_goal_skip.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_longlong]
    L.shim_same_t_bound.restype = ctypes.c_int
    L.shim_same_t_bound.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_int, ctypes.c_int]
    L.shim_gate_info.restype = ctypes.c_int
    L.shim_gate_info.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                 ctypes.c_int]
    return L


def _local(name):
    if name in EXTRA:
        gk, loc = EXTRA[name]
        nq = 2
    else:
        gk, loc = W._GATES[name]
        nq = {"cx": 2, "cz": 2, "ccx": 3}.get(name, 1)
    Ld = 1 << nq
    M = [[W.ZERO] * Ld for _ in range(Ld)]
    for (r, c), p in loc.items():
        M[r][c] = W._wpow(p)
    if name == "ch":   # control 0 (local bit 0 clear): identity times sqrt2 = w - w^3, scaled back by gk = 1
        for r in (0, 2):
            M[r][r] = (0, 1, 0, -1)
    return nq, gk, M


def _conj(z):
    return (z[0], -z[3], -z[2], -z[1])


def _spec(nq, gk, M):
    Ld = 1 << nq
    fwd = " ".join(" ".join(map(str, M[r][c])) for r in range(Ld) for c in range(Ld))
    inv = " ".join(" ".join(map(str, _conj(M[c][r]))) for r in range(Ld) for c in range(Ld))
    N = Ld
    ident = "".join(" ".join(" ".join(map(str, W.ONE if i == j else W.ZERO)) for j in range(N)) + "\n"
                    for i in range(N))
    qs = " ".join(str(q) for q in range(nq))
    cuts = "".join(f"1 {q}\n" for q in range(nq))
    return (f"cheese-spec 1\nn {nq}\nmode unitary\nphase exact\ndepth 1\nstandard 1\ngates 1\n"
            f"L_g_{'_'.join(str(q) for q in range(nq))} {nq} {qs} {gk} {fwd} {gk} {inv}\n"
            f"bad2 0\nbad3 0\ntarget 0\n{ident}cuts {nq}\n{cuts}cxtable 0\nend\n")


@pytest.mark.parametrize("name", ["x", "z", "h", "s", "sdg", "t", "tdg", "cx", "cz", "ccx", "cs", "ct", "ch"])
def test_gate_analysis_matches_the_oracle(shim, oracle_lib, name):
    nq, gk, M = _local(name)
    spec = _spec(nq, gk, M).encode()
    out = (ctypes.c_int * 8)()
    nc = shim.shim_gate_info(spec, len(spec), 0, out, 8)
    assert nc == nq
    flat = [c for row in M for z in row for c in z]
    buf = (ctypes.c_longlong * len(flat))(*flat)
    sde = oracle_lib.orc_bound(nq, gk, buf, 0, 0)         # T bound of the gate's matrix = its PTM exponent
    assert out[1] == sde, (name, out[1], sde)
    assert out[0] == (1 if sde == 0 else 0)
    for q in range(nq):
        assert out[2 + q] == oracle_lib.orc_bound(nq, gk, buf, 4, 1 << q), (name, q)


def test_known_exponents_and_ranks(shim):
    """closed forms: Clifford gates 0, T 1, CCX and CS 2 (PTM entries 1/2); CX and CCX have Schmidt rank 2 across
    any cut they straddle"""
    want = {"h": 0, "s": 0, "cx": 0, "cz": 0, "t": 1, "tdg": 1, "ccx": 2, "cs": 2}
    for name, e in want.items():
        nq, gk, M = _local(name)
        spec = _spec(nq, gk, M).encode()
        out = (ctypes.c_int * 8)()
        assert shim.shim_gate_info(spec, len(spec), 0, out, 8) == nq
        assert out[1] == e, name
        if name in ("cx", "ccx"):
            assert all(out[2 + q] == 2 for q in range(nq)), name


@pytest.mark.parametrize("n", [2, 3])
def test_goal_preimage_is_the_candidate_itself(shim, n):
    """the residual from which one candidate reaches the goal (the identity) is that candidate's own matrix:
    closed form, built here from the gate tables of the workload generator, not from the spec's matrices.  A
    residual one gate short of the identity is what flags a parent whose level ends the search."""
    skel = W.skeleton(n)
    spec = W.spec_from_skeleton(skel, W.unitary(n, []), 5).encode()
    names = [ln.split()[0] for ln in skel.splitlines() if ln.startswith("L_")]
    N = 1 << n
    buf = (ctypes.c_longlong * (4 * N * N))()
    for g, name in enumerate(names):
        parts = name.split("_")
        gate, qs = parts[1], [int(q) for q in parts[2:]]
        k = shim.shim_goal_preimage(spec, len(spec), g, buf, len(buf))
        want_k, want = W.unitary(n, [(gate, qs)])
        assert k == want_k, name
        got = [tuple(buf[4 * i:4 * i + 4]) for i in range(N * N)]
        assert got == [x for row in want for x in row], name


@pytest.mark.parametrize("name", ["cs-d4", "toffoli-d9", "rus13-unitary-d12"])
def test_found_level_skip_needs_every_preimage(shim, name):
    """the search skips goal-free parents at a found level only when every candidate has a goal pre-image; goal
    entries scaled beyond int32 leave candidates without one, which turns the skip off"""
    spec = read_spec(name).encode()
    assert shim.shim_goal_skip(spec, len(spec), 1) == 1
    assert shim.shim_goal_skip(spec, len(spec), 1 << 40) == 0


def test_t_bound_classes_share_the_oracle_bound(shim, oracle_lib):
    """candidates a, b with a b^-1 Clifford give children of equal T bound from any residual, so the expansion
    computes one T transform per class and parent.  On the Toffoli candidate set the host pairs exactly T and
    T-dagger on each qubit among its non-Clifford candidates (Clifford children inherit their parent's bound
    anyway); the oracle's T bound (the definitional transfer-matrix exponent) of S a^-1 and S b^-1 agrees
    on random residuals S for every pair, and differs for T on two different qubits (a pairing the rule rejects)"""
    import random
    spec = read_spec("toffoli-d12").encode()
    names = [ln.split()[0] for ln in read_spec("toffoli-d12").split("gates ")[1].splitlines()[1:28]]
    G = len(names)
    info = (ctypes.c_int * 8)()
    nonclifford = []
    for g in range(G):
        shim.shim_gate_info(spec, len(spec), g, info, 8)
        if not info[0]:
            nonclifford.append(g)
    pairs = {(a, b) for a in nonclifford for b in nonclifford
             if a < b and shim.shim_same_t_bound(spec, len(spec), a, b) == 1}
    want = {(names.index(f"L_t_{q}"), names.index(f"L_tdg_{q}")) for q in range(3)}
    assert pairs == want
    rng = random.Random(5)
    N, E = 8, 64

    def tbound(seq):
        buf = (ctypes.c_longlong * (4 * E))()
        arr = (ctypes.c_int * max(1, len(seq)))(*seq)
        k = oracle_lib.orc_residual(spec, len(spec), arr, len(seq), buf)
        assert k >= 0
        return oracle_lib.orc_bound(3, k, buf, 1, 0)

    differ = 0
    for _ in range(40):
        prefix = [rng.randrange(G) for _ in range(rng.randrange(0, 6))]
        for a, b in want:
            assert tbound(prefix + [a]) == tbound(prefix + [b])
        differ += tbound(prefix + [names.index("L_t_0")]) != tbound(prefix + [names.index("L_t_1")])
    assert differ > 0


def test_zdown_halves_in_32_bits():
    """the device's division by sqrt2 (dev.h zdown) forms (a + c) / 2 and (c - a) / 2 of int32 coefficients of equal
    parity as (a >> 1) + (c >> 1) + (a & 1) and (c >> 1) - (a >> 1), with no 64-bit sum: equal to the exact halves on
    every pair of extreme values and on random ones, and within int32"""
    import random
    lo, hi = -(1 << 31), (1 << 31) - 1
    edge = [lo, lo + 1, lo + 2, -3, -2, -1, 0, 1, 2, 3, hi - 2, hi - 1, hi]
    rnd = random.Random(11)
    vals = edge + [rnd.randint(lo, hi) >> rnd.randint(0, 31) for _ in range(4000)]
    n = 0
    for a in vals[:200] if len(vals) > 200 else vals:
        for c in vals:
            if (a ^ c) & 1:
                continue
            s, d = (a >> 1) + (c >> 1) + (a & 1), (c >> 1) - (a >> 1)
            assert s == (a + c) // 2 and d == (c - a) // 2 and lo <= s <= hi and lo <= d <= hi, (a, c)
            n += 1
    assert n > 100000
