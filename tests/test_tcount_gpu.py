"""The HIP T-count path (coset table in HBM, batched searches) against the reference's own logs and results
(benchmarks/optimal/*-tcount.json), the published table sizes (n=2 K=6: 1,728,736; n=3 K=4: 5,605,741), the
T-count oracle on seeded random targets, and exact verification of every decomposition."""
import ctypes
import json
import os

import pytest

from conftest import GOLDEN, read_tspec, strip_seconds

from cheese_amd import tcount as T
from cheese_amd import workloads as W

pytestmark = pytest.mark.gpu

TC = os.path.join(GOLDEN, "tcount")
NAMES2 = ["cs", "qft_2"]
NAMES3 = ["relative-phase-toffoli", "toffoli", "ccz", "peres", "fredkin"]


def _golden(name):
    return strip_seconds(open(os.path.join(TC, name)).read())


@pytest.fixture(scope="module")
def table2():
    t = T.CosetTable(read_tspec("cs"), K=6)
    yield t
    t.close()


@pytest.fixture(scope="module")
def table3():
    t = T.CosetTable(read_tspec("toffoli"), K=4)
    yield t
    t.close()


def _unitary(name):
    return W.parse_qasm(open(os.path.join(TC, name + ".qasm")).read())


def test_two_qubit_table(table2):
    assert table2.size == 1728736
    assert strip_seconds("\n".join(table2.table_lines())) == _golden("table-n2-k6.oracle.log")


def test_three_qubit_table(table3):
    assert table3.size == 5605741
    assert strip_seconds("\n".join(table3.table_lines())) == _golden("table-n3-k4.oracle.log")


@pytest.mark.parametrize("name", NAMES2)
def test_two_qubit_searches_match_reference(table2, name):
    r = table2.search([read_tspec(name)], maxt=16)[0]
    assert r["rc"] == 0
    assert strip_seconds("\n".join(r["lines"])) == _golden(name + ".json")


def test_three_qubit_batch_matches_reference(table3):
    """all five reference targets in one batched search: every log and result as the reference printed it,
    and the same results as one target at a time"""
    specs = [read_tspec(n) for n in NAMES3]
    batch = table3.search(specs, maxt=16)
    for name, r in zip(NAMES3, batch):
        assert r["rc"] == 0, name
        assert strip_seconds("\n".join(r["lines"])) == _golden(name + ".json"), name
        n, circ = _unitary(name)
        assert T.verify(n, W.unitary(n, circ), [list(x) for x in r["rotations"]])
    for name, r in zip(NAMES3[:2], batch[:2]):
        alone = table3.search([read_tspec(name)], maxt=16)[0]
        assert strip_seconds("\n".join(alone["lines"])) == strip_seconds("\n".join(r["lines"]))


def test_budget_too_small_and_clifford_target(table3):
    """maxt below the T-count gives the tcount_above line (exit 1); a Clifford target has T-count 0"""
    r = table3.search([read_tspec("toffoli")], maxt=6)[0]
    assert r["rc"] == 1
    assert json.loads(r["lines"][-1])["tcount_above"] == 6
    n = 3
    U = W.unitary(n, [("h", (0,)), ("cx", (0, 1)), ("s", (2,))])
    r = table3.search([T.tspec_text(n, U)], maxt=4)[0]
    assert r["rc"] == 0 and r["tcount"] == 0 and r["split"] == [0, 0] and r["levels"] == []


def test_bad_specs(table3):
    good = read_tspec("toffoli")
    res = table3.search(["not a spec\n", read_tspec("cs"), good], maxt=16)
    assert res[0]["rc"] == 2
    assert res[1]["rc"] == 3          # two-qubit spec against a three-qubit table
    assert res[2]["rc"] == 0 and res[2]["tcount"] == 7


def _clifford_words(r):
    return T.random_clifford_word(r, n=2, lmin=4, lmax=12)


def test_random_two_qubit_targets_match_oracle(table2, tc_oracle):
    """24 seeded random two-qubit targets in one batch: logs and results equal the oracle's"""
    specs = []
    for seed in range(24):
        k, circ = T.random_target(1000 + seed, n=2, kmin=1, kmax=8, clifford_words=_clifford_words)
        specs.append(T.tspec_text(2, W.unitary(2, circ)))
    got = table2.search(specs, maxt=12)
    buf = ctypes.create_string_buffer(1 << 16)
    s0 = specs[0].encode()
    h = tc_oracle.orc_tc_table(s0, len(s0), 6, 0, buf, len(buf))
    try:
        for spec, r in zip(specs, got):
            raw = spec.encode()
            rot = (ctypes.c_int * 64)()
            rc = tc_oracle.orc_tc_search(h, raw, len(raw), 12, buf, len(buf), rot)
            assert rc == r["rc"]
            assert strip_seconds("\n".join(r["lines"])) == strip_seconds(buf.value.decode())
    finally:
        tc_oracle.orc_tc_free(h)


def test_random_three_qubit_batch_verifies(table3):
    """16 seeded random three-qubit targets (T-count <= k by construction) in one batch: each decomposition is
    exact, no longer than k, and no single rotation can be dropped"""
    cases = []
    for seed in range(16):
        k, circ = T.random_target(2000 + seed, n=3, kmin=2, kmax=7,
                                  clifford_words=lambda r: T.random_clifford_word(r, n=3, lmin=4, lmax=12))
        cases.append((k, W.unitary(3, circ)))
    got = table3.search([T.tspec_text(3, U) for _, U in cases], maxt=16)
    for (k, U), r in zip(cases, got):
        assert r["rc"] == 0
        assert r["tcount"] <= k
        rots = [list(x) for x in r["rotations"]]
        assert T.verify(3, U, rots)
        for drop in range(len(rots)):
            assert not T.verify(3, U, rots[:drop] + rots[drop + 1:])


def test_batched_targets_from_clifford_samples(table3):
    """the batched workload at its full size, 256 targets (Cliffords sampled uniformly from the Clifford table,
    realised by their optimal words) in one call: seeded sampling is reproducible, every decomposition is exact and
    within k"""
    from cheese_amd import clifford
    cl = clifford.CliffordTable(3)
    try:
        a, b = cl.sample(5, 100), cl.sample(5, 100)
        assert a == b and len(set(a)) > 90
        assert all(w is not None for w in cl.readback(a[:20]))
        cases = T.c3_targets(cl, 256, seed0=77)
    finally:
        cl.close()
    got = table3.search([T.tspec_text(3, W.unitary(3, c)) for _, c in cases], maxt=16)
    for (k, circ), r in zip(cases, got):
        assert r["rc"] == 0 and r["tcount"] <= k
        assert T.verify(3, W.unitary(3, circ), [list(x) for x in r["rotations"]])


def test_tcount_command_line_matches_reference():
    """`tcount toffoli.tspec --table 4`: the table lines, search lines and result line of the reference"""
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "cheese_amd", "bin", "tcount")
    p = subprocess.run([exe, os.path.join(GOLDEN, "tspec", "toffoli.tspec"), "--table", "4"], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0
    lines = p.stdout.strip().splitlines()
    assert strip_seconds("\n".join(lines[:4])) == _golden("table-n3-k4.oracle.log")
    assert strip_seconds("\n".join(lines[4:])) == _golden("toffoli.json")
