"""The outer budget loop (cheese_amd/run.py, the reference's run.sh:1-64): QASM targets become the reference's own
.spec text (byte-identical to the golden specs its writer produced, through cheese_amd/specgen.py), found circuits are written as OpenQASM that
reads back to the same gates, and the loop's logs and optimum match the reference's golden logs on the GPU."""
import json
import os

import pytest

from conftest import GOLDEN, manifest, read_log, read_spec, strip_seconds

from cheese_amd import run as R
from cheese_amd import workloads as W


def _qasm(name):
    return os.path.join(GOLDEN, "qasm", name + ".qasm")


@pytest.mark.parametrize("case", sorted(manifest()))
def test_spec_of_a_qasm_target_is_the_references(case):
    m = manifest()[case]
    if "qasm" not in m or m.get("flags"):
        pytest.skip("no QASM source, or a spec written with extra writer options")
    _, _, _, base = R.target_spec(_qasm(m["qasm"][:-5]))
    assert R.with_depth(base, m["depth"]) == read_spec(case)


def test_qasm_writer_reads_back():
    names = ["L_h_2", "L_cx_1_2", "L_tdg_2", "L_cx_0_2", "L_t_1", "L_s_0"]
    n, circ = W.parse_qasm(R.qasm_text(3, names))
    assert n == 3 and circ == R.gates_of(names)


def test_unsupported_targets_and_missing_files(tmp_path):
    q = tmp_path / "four.qasm"
    q.write_text('OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[4];\nh q[3];\ncx q[0],q[3];\n')
    msgs = []
    assert R.run(str(q), results=str(tmp_path / "r"), echo=msgs.append) == 3
    assert "one to three" in msgs[-1]
    assert R.run(str(tmp_path / "none.qasm"), results=str(tmp_path / "r"), echo=msgs.append) == 2


def _check_log(path, case):
    with open(path) as f:
        got = strip_seconds(f.read())
    ref = strip_seconds(read_log(case))
    assert got == ref, case


@pytest.mark.gpu
@pytest.mark.parametrize("target,first,proof,found", [("cs", 1, "cs-d4", "cs-d5"), ("qft_2", 6, "qft_2-d6", "qft_2-d7")])
def test_budget_loop_finds_the_optimum(tmp_path, target, first, proof, found):
    """budgets below the optimum end in proofs, the optimum's budget finds the reference's circuit; each budget's
    log equals the reference's golden log; the written QASM has the target's exact unitary"""
    res = tmp_path / "results"
    msgs = []
    rc = R.run(_qasm(target), first=first, results=str(res), echo=msgs.append)
    assert rc == 0, msgs
    dp, df = manifest()[proof]["depth"], manifest()[found]["depth"]
    _check_log(res / f"{target}-d{dp}.log", proof)
    _check_log(res / f"{target}-d{df}.log", found)
    assert not (res / f"{target}-d{df + 1}.log").exists()
    assert any(m.startswith(f"optimum: ") for m in msgs)
    with open(_qasm(target)) as f:
        n, circ = W.parse_qasm(f.read())
    with open(res / f"{target}-opt.qasm") as f:
        n2, got = W.parse_qasm(f.read())
    assert n2 == n and W.unitary(n, got) == W.unitary(n, circ)
    last = json.loads(read_log(found).strip().splitlines()[-1])
    assert len(got) == last["t"]


@pytest.mark.gpu
def test_single_budget_proof(tmp_path):
    msgs = []
    assert R.run(_qasm("cs"), only=4, results=str(tmp_path), echo=msgs.append) == 1
    assert msgs[-1].startswith("budget 4: no circuit of <= 4 gates (proof")
    _check_log(tmp_path / "cs-d4.log", "cs-d4")
