"""cliffopt, the T-placement beam search over the Clifford table (src/clifford.cu:168-206, :285-385): the oracle's
CPU restatement is pinned by exact equivalence (the output circuit equals the job's circuit up to a global phase,
with the same T gates in number) and by minimality of every output segment; the HIP path (chz_cliff_opt) equals
the oracle's output line for line and keeps the same invariants at the reference's default beam."""
import ctypes
import os
import subprocess

import pytest

from conftest import ROOT

from cheese_amd import clifford as CL
from cheese_amd import workloads as W


def _wmul(p, x):
    return W._zmul(W._wpow(p), x)


def equal_up_to_phase(U, V):
    (ka, A), (kb, B) = U, V
    if ka != kb:
        return False
    for p in range(8):
        if all(_wmul(p, A[r][c]) == B[r][c] for r in range(len(A)) for c in range(len(A))):
            return True
    return False


@pytest.fixture(scope="module")
def oracle2():
    L = ctypes.CDLL(os.path.join(ROOT, "oracle", "libclifford_oracle.so"))
    L.orc_cliff_build.restype = ctypes.c_void_p
    L.orc_cliff_build.argtypes = [ctypes.c_int, ctypes.c_int]
    L.orc_cliff_opt.restype = ctypes.c_int
    L.orc_cliff_opt.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_long, ctypes.c_int, ctypes.c_char_p,
                                ctypes.c_long]
    L.orc_cliff_free.argtypes = [ctypes.c_void_p]
    L.orc_cliff_find.restype = ctypes.c_int
    L.orc_cliff_find.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    L.orc_cliff_word.restype = ctypes.c_uint64
    L.orc_cliff_word.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    h = L.orc_cliff_build(2, 0)
    yield L, h
    L.orc_cliff_free(h)


def oracle_opt(L, h, job, beam=16):
    buf = ctypes.create_string_buffer(1 << 20)
    raw = job.encode()
    rc = L.orc_cliff_opt(h, raw, len(raw), beam, buf, len(buf))
    assert rc == 0
    return buf.value.decode()


def _segments(circ):
    """the Clifford runs between T gates"""
    segs, cur = [], []
    for g, qs in circ:
        if g in ("t", "tdg"):
            segs.append(cur)
            cur = []
        else:
            cur.append((g, qs))
    segs.append(cur)
    return segs


@pytest.mark.parametrize("seed", range(6))
def test_oracle_output_is_equivalent_and_minimal(oracle2, seed):
    L, h = oracle2
    job, circ = CL.random_job(seed, n=2, t=3 + seed % 3)
    head, out = CL.parse_opt_output(oracle_opt(L, h, job))
    assert head["t_gates"] == sum(g in ("t", "tdg") for g, _ in circ) == sum(g in ("t", "tdg") for g, _ in out)
    assert head["clifford_gates"] == sum(g not in ("t", "tdg") for g, _ in out)
    assert head["clifford_gates_before"] == sum(g not in ("t", "tdg") for g, _ in circ)
    assert head["clifford_gates"] <= head["clifford_gates_before"]
    assert equal_up_to_phase(W.unitary(2, circ), W.unitary(2, out))
    # every output segment is an optimal word: its length is the table's distance of its tableau
    for seg in _segments(out):
        idx = [_gate_index(g, qs) for g, qs in seg]
        arr = (ctypes.c_int * max(1, len(idx)))(*idx)
        assert L.orc_cliff_find(h, L.orc_cliff_word(h, arr, len(idx))) == len(seg)


def _gate_index(g, qs, n=2):
    """index into the table's gate set: H, S, Sdg, X, Z per qubit, then CX on ordered pairs"""
    if g == "cx":
        pairs = [(a, b) for a in range(n) for b in range(n) if a != b]
        return 5 * n + pairs.index(tuple(qs))
    return 5 * qs[0] + ("h", "s", "sdg", "x", "z").index(g)


def test_job_gate_indices_follow_the_gate_set():
    assert [_gate_index("h", (0,)), _gate_index("z", (1,)), _gate_index("cx", (0, 1)), _gate_index("cx", (1, 0))] == \
        [0, 9, 10, 11]


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(6))
def test_hip_equals_oracle_two_qubits(oracle2, seed):
    L, h = oracle2
    job, _ = CL.random_job(seed, n=2, t=3 + seed % 3)
    print(f"cliffopt n=2 seed {seed}", flush=True)
    want = oracle_opt(L, h, job).strip().splitlines()
    t = CL.CliffordTable(2)
    try:
        got = t.optimize(job).strip().splitlines()
    finally:
        t.close()
    import json
    g0 = json.loads(got[0])
    g0.pop("seconds")
    assert g0 == json.loads(want[0])
    assert got[1:] == want[1:]


@pytest.mark.gpu
def test_hip_three_qubits_equivalent_and_shorter():
    t = CL.CliffordTable(3)
    try:
        for seed in range(4):
            job, circ = CL.random_job(100 + seed, n=3, t=4, lmin=6, lmax=14)
            print(f"cliffopt n=3 seed {100 + seed}", flush=True)
            head, out = CL.parse_opt_output(t.optimize(job, beam=16))
            assert head["t_gates"] == 4 == sum(g in ("t", "tdg") for g, _ in out)
            assert head["clifford_gates"] <= head["clifford_gates_before"]
            assert equal_up_to_phase(W.unitary(3, circ), W.unitary(3, out))
    finally:
        t.close()


@pytest.mark.gpu
def test_cliffopt_command_line(tmp_path):
    job, circ = CL.random_job(7, n=2, t=3)
    p = tmp_path / "job.txt"
    p.write_text(job)
    r = subprocess.run([os.path.join(ROOT, "cheese_amd", "bin", "cliffopt"), str(p), "--beam", "8"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    head, out = CL.parse_opt_output(r.stdout)
    assert equal_up_to_phase(W.unitary(2, circ), W.unitary(2, out))
