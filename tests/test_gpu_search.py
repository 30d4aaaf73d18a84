"""Parity of the HIP gate-count search with the reference: per-level lines and result lines identical to the
reference CPU build's golden logs, every emitted circuit exactly re-verified.  Needs an MI355X."""
import pytest

from conftest import flags_to_kwargs, golden_cases, manifest, read_log, read_spec, strip_seconds

import cheese_amd
from cheese_amd import workloads

pytestmark = pytest.mark.gpu


def _lines(out):
    lines = [cheese_amd.level_line(l) for l in out["levels"]]
    if out["rc"] == 0 or (out["rc"] == 1 and out.get("split")):
        lines.append(cheese_amd.result_line(out))
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("name", golden_cases())
def test_hip_matches_reference_log(name):
    """every golden log, the Toffoli optimum-15 proofs (budgets 13 and 14) and the random circuits the reference
    needed up to 24 minutes of CPU for included"""
    m = manifest()[name]
    print(f"case {name}", flush=True)
    out = cheese_amd.search(read_spec(name), **flags_to_kwargs(m["flags"]))
    assert out["rc"] == m["rc"], out.get("error")
    assert strip_seconds(_lines(out)) == strip_seconds(read_log(name))
    if out["rc"] == 0:
        assert out["verified"]


def test_chunking_leaves_levels_unchanged():
    """--chunk only changes how a level is split into passes (src/cheese.cu:1264)"""
    spec = read_spec("toffoli-d10")
    a = cheese_amd.search(spec, chunk=37)
    assert strip_seconds(_lines(a)) == strip_seconds(read_log("toffoli-d10"))


@pytest.mark.parametrize("seed", range(32))
def test_random_circuit_found_and_verified(seed):
    """C2's seeded random 3-qubit Clifford+T circuits at budget = length: a circuit of at most that length,
    exactly equal to the target; identical to the reference log where the reference finished"""
    name = f"random3-s{seed}"
    m = manifest().get(name)
    circ, spec = workloads.random_spec(seed)
    print(f"case {name} length {len(circ)}", flush=True)
    out = cheese_amd.search(spec)
    assert out["rc"] == 0
    assert out["verified"]
    assert len(out["circuit"]) <= len(circ)
    if m and m.get("rc") is not None:
        assert strip_seconds(_lines(out)) == strip_seconds(read_log(name))


def test_overflow_exit_4():
    """a target whose coefficients leave int32 aborts with exit 4 (src/cheese.cu:1797)"""
    spec = read_spec("cs-d4")
    lines = spec.splitlines()
    i = lines.index([l for l in lines if l.startswith("target ")][0])
    row = lines[i + 1].split()
    row[0] = str(1 << 40)
    lines[i + 1] = " ".join(row)
    out = cheese_amd.search("\n".join(lines) + "\n")
    assert out["rc"] == 4


def test_depth_zero_identity():
    """a target equal to the goal prints {"found": 0, "circuit": []} (src/cheese.cu:1794-1801)"""
    skel = workloads.skeleton(2)
    out = cheese_amd.search(skel)
    assert out["rc"] == 0 and out["circuit"] == []


def test_expansion_stage_counts_add_up():
    """the expansion's stage counters (the factors of its work model) against identities of the level lines: every
    child not cut by a window is formed, every survivor of a level that keeps its survivors gets a class key, T-bound
    evaluations and cut ranks happen only on formed children"""
    spec = read_spec("toffoli-d10")
    s = cheese_amd.Searcher(spec)
    try:
        out = s.run(stats=True)
    finally:
        s.close()
    st, lv = out["expand_stages"], out["levels"]
    depth = int(spec.split("\ndepth ")[1].split()[0])
    assert out["rc"] == 1 and lv
    assert st["residuals"] == sum(l["prefixes"] - l["cut_w3"] for l in lv)
    # a level's survivors get keys while its search goes on (t < depth)
    assert st["keys"] == sum(l["survivors"] for l in lv if l["t"] < depth)
    assert st["structural"] <= st["residuals"]
    assert 0 < st["t_bounds"] < st["structural"]
    assert 0 < st["cut_ranks"] <= 3 * st["structural"]
    # a T-bound evaluation runs at most both flip-pair passes, and the lower bounds settle most children early
    assert 0 < st["t_passes"] < 2 * st["t_bounds"]


def test_table_size_within_seven_gates():
    """docs/PLAN.md:87: the 3-qubit Clifford+T distance table of K = 7 holds 2.65 M classes (the identity and each
    table level's new classes)"""
    out = cheese_amd.search(read_spec("toffoli-d9"), table=7)
    lines = [l for l in out["levels"] if l["dir"] == 1]
    assert len(lines) == 7
    total = 1 + sum(l["distinct"] for l in lines)
    print(f"table K=7: {total} classes", flush=True)
    assert abs(total - 2.65e6) < 0.005e6
    assert out["rc"] == 1


def test_toffoli_optimum_fifteen_with_table_eight():
    """docs/PLAN.md:87-91: K = 8 holds 17.9 M classes; with --table 8 budget 14 is proved impossible and budget 15
    finds a circuit, so Toffoli's optimal gate count is 15; the circuit is checked exactly against the target"""
    spec14 = read_spec("toffoli-d14")
    out = cheese_amd.search(spec14, table=8)
    total = 1 + sum(l["distinct"] for l in out["levels"] if l["dir"] == 1)
    print(f"table K=8: {total} classes, budget 14 rc {out['rc']}", flush=True)
    assert abs(total - 17.9e6) < 0.05e6
    assert out["rc"] == 1 and out["split"] == ("table", [-1, 8])
    out = cheese_amd.search(spec14.replace("\ndepth 14\n", "\ndepth 15\n"), table=8)
    assert out["rc"] == 0 and out["verified"] and len(out["circuit"]) == 15


@pytest.mark.parametrize("seed", [0, 3, 5, 9, 17, 21])
@pytest.mark.parametrize("mode", ["table", "mitm"])
def test_table_and_join_find_circuits_of_the_plain_length(seed, mode):
    """seeded random circuits at budget = length: --table 5 and --mitm find circuits of the length the plain search
    finds, each exactly equal to the target"""
    circ, spec = workloads.random_spec(seed)
    print(f"case random3-s{seed} {mode}", flush=True)
    plain = cheese_amd.search(spec)
    other = cheese_amd.search(spec, **({"table": 5} if mode == "table" else {"mitm": True}))
    assert plain["rc"] == other["rc"] == 0
    assert other["verified"]
    assert len(other["circuit"]) == len(plain["circuit"])


def _cuts(spec):
    """(qubit mask, radix) of each cut of a .spec: its `cuts` section lists the size and the qubits of one side"""
    lines = spec.splitlines()
    i = next(j for j, l in enumerate(lines) if l.startswith("cuts "))
    out = []
    for l in lines[i + 1:i + 1 + int(lines[i].split()[1])]:
        f = [int(x) for x in l.split()]
        out.append((sum(1 << q for q in f[1:1 + f[0]]), 2 * f[0] + 1))
    return out


@pytest.mark.parametrize("name", ["toffoli-d12", "random3-s5"])
def test_frontier_bound_words_equal_the_oracle(oracle_lib, name):
    """the bound word every frontier node carries holds the oracle's exact values for the node's own residual (the
    residual of its gate path from the target): its T bound (src/cheese.cu:397-449) wherever the word says it is
    known, whether inherited from the parent, shared within a T-bound class or settled by the staged transform, and
    each cut's need, the clipped ceil log2 of the F_p operator-Schmidt rank (:270-303, :585-597), wherever the
    needs are known; sampled after every level"""
    import ctypes
    L = oracle_lib
    spec = read_spec(name)
    raw = spec.encode()
    cuts = _cuts(spec)
    b = cheese_amd.Batch(spec)
    checked_t = checked_c = 0
    try:
        assert b.add(spec, 0) == 0
        for level in range(1, 16):
            b.events()
            if not b.live:
                break
            assert b.step() == 0
            if not b.live or b.nodes == 0:
                break
            for tag, path, w in b.frontier(120, max(1, b.nodes // 120)):
                assert tag == 0
                e = (ctypes.c_longlong * (4 * 64))()
                arr = (ctypes.c_int * max(1, len(path)))(*path)
                k = L.orc_residual(raw, len(raw), arr, len(path), e)
                assert k >= 0, path
                if w & 0xff != 0xff:
                    assert w & 0xff == L.orc_bound(3, k, e, 0, 0), (level, path, w)
                    checked_t += 1
                if w >> 31:
                    for c, (mask, rad) in enumerate(cuts):
                        rank, need = L.orc_bound(3, k, e, 4, mask), 0
                        while (1 << need) < rank:
                            need += 1
                        assert (w >> (8 + 3 * c)) & 7 == min(need, rad - 1), (level, path, c, w)
                    checked_c += 1
    finally:
        b.close()
    print(f"case {name}: {checked_t} T bounds and {checked_c} cut-need sets equal to the oracle's", flush=True)
    assert checked_t >= 200 and checked_c >= 200


def _one_qubit_case(name):
    """(circuit, spec at budget = length): seeded random 1-qubit Clifford+T words, or (H T)^k, whose shortest
    circuit is the word itself"""
    from cheese_amd import specgen
    kind, v = name.split("-")
    if kind == "ht":
        circ = [("h", (0,)), ("t", (0,))] * int(v)
    else:
        circ = workloads.random_circuit(int(v), 1, 8, 14, gates=("x", "z", "h", "s", "sdg", "t", "tdg"))
    return circ, specgen.spec_text(1, circ, len(circ))


@pytest.mark.parametrize("name", ["rand-2", "rand-3", "rand-5", "ht-5", "ht-8"])
def test_one_qubit_search_matches_the_oracle(oracle_lib, name):
    """the n = 1 instantiation of the level kernels (lane groups of 4 entries, 16 groups per wave, per-group status
    tallies): level lines and result line equal the oracle's (src/cheese.cu:1249-1377) at budget = length, and at
    one gate below the circuit found, where the search proves the budget impossible"""
    from conftest import oracle_run
    circ, spec = _one_qubit_case(name)
    print(f"case n1 {name} length {len(circ)}", flush=True)
    out = cheese_amd.search(spec)
    rc, ref = oracle_run(oracle_lib, spec)
    assert out["rc"] == rc == 0 and out["verified"]
    assert strip_seconds(_lines(out)) == strip_seconds(ref)
    d = len(out["circuit"]) - 1
    assert d >= 1
    spec_d = spec.replace(f"\ndepth {len(circ)}\n", f"\ndepth {d}\n")
    out = cheese_amd.search(spec_d)
    rc, ref = oracle_run(oracle_lib, spec_d)
    assert out["rc"] == rc == 1
    assert strip_seconds(_lines(out)) == strip_seconds(ref)
