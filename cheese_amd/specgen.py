"""The .spec writer for any unitary target over the gate vocabulary (python/cheese/spec.py:127-226, :356-415):
candidates, the W2/W3 window rules found by exact products, the CX cut table, the bound flags and the spec text,
so that targets outside the default candidate set need not go through the reference's Python.

  python -m cheese_amd.specgen TARGET.qasm DEPTH > target.spec

The gate set is the default one (cx h s sdg t tdg x z) plus the target's own gates, sorted by name
(spec.py:357).  State targets (--ancilla / --inputs) are not written: the search here answers them with exit 3.
"""
from __future__ import annotations

import io
import itertools
import sys

from . import workloads as W

DAGGER = {"s": "sdg", "sdg": "s", "t": "tdg", "tdg": "t"}       # every other gate is its own inverse
STANDARD = frozenset({"x", "y", "z", "s", "sdg", "t", "tdg", "cx", "cz", "h"})
# Pauli-transfer-matrix exponent of each non-Clifford gate of the vocabulary: how far one gate can move the T bound
# (spec.py:317-334; the host's gate_ptm_sde computes the same numbers, checked in tests/test_specgen.py)
PTM_SDE = {"t": 1, "tdg": 1, "ccx": 2}


def arity(g: str) -> int:
    return {"cx": 2, "cz": 2, "ccx": 3}.get(g, 1)


def candidates(n: int, gates):
    """(name, gate, qubits) in the reference's order: gate by gate as given, then placements (spec.py:127-136)"""
    out = []
    for g in gates:
        a = arity(g)
        if a == 1:
            places = [(q,) for q in range(n)]
        elif a == 2:
            places = [(c, t) for c in range(n) for t in range(n) if c != t]
        else:                                     # ccx: its two controls unordered
            places = [qs for qs in itertools.permutations(range(n), a) if qs[0] < qs[1]]
        out.extend((f"L_{g}_" + "_".join(str(q) for q in qs), g, qs) for qs in places)
    return out


def _acts_trivially(M, bit):
    """M = M' (x) I on the local qubit `bit`: every entry equals the matching entry of the block where that
    qubit's row and column bits are 0 when the two bits agree, and is zero when they differ"""
    N = len(M)
    for r in range(N):
        for c in range(N):
            want = W.ZERO if (r ^ c) & bit else M[r & ~bit][c & ~bit]
            if M[r][c] != want:
                return False
    return True


def product_key(cs, seq):
    """exact identity of the product of candidates seq (first applied first): the qubits it acts on
    nontrivially, its sqrt2 exponent and its normalised entries on those qubits"""
    support = sorted({q for g in seq for q in cs[g][2]})
    local = [(cs[g][1], tuple(support.index(q) for q in cs[g][2])) for g in seq]
    k, M = W.unitary(len(support), local)
    idle = [i for i in range(len(support)) if _acts_trivially(M, 1 << i)]
    if idle:
        drop = sum(1 << i for i in idle)
        rows = [r for r in range(len(M)) if not r & drop]      # the block on the remaining qubits
        k, M = W.normalise(k, [[M[r][c] for c in rows] for r in rows])
        support = [q for i, q in enumerate(support) if i not in idle]
    return tuple(support), k, tuple(map(tuple, M))


def _cut_windows(windows, keys, reducible):
    """the windows (given in enumeration order) that W2/W3 cut: a window whose product equals a shorter one
    (`reducible`), and every window of an equal-product class except its earliest in enumeration order"""
    earliest = {}
    for i, key in enumerate(keys):
        earliest.setdefault(key, i)
    return [w for i, (w, key) in enumerate(zip(windows, keys)) if key in reducible or earliest[key] != i]


def window_rules(cs, use3: bool = True, max_triples: int = 300_000):
    """W2 pairs and W3 triples of candidate indices.  Pairs are enumerated lexicographically.  Triples extend
    the allowed pairs (f, g), taken in sorted order, by every h such that (g, h) is allowed, the h in the
    iteration order of the allowed-pair set built as the set difference of all pairs and the cut pairs; which
    member of an equal-product class is kept depends on that order, so it is the writer's contract
    (python/cheese/spec.py:165-185).  No triples are cut when there are more than max_triples of them."""
    G = len(cs)
    lower = {product_key(cs, ())} | {product_key(cs, (g,)) for g in range(G)}
    pairs = list(itertools.product(range(G), repeat=2))
    pair_keys = [product_key(cs, p) for p in pairs]
    bad2 = _cut_windows(pairs, pair_keys, lower)
    allowed = set(pairs) - set(bad2)
    after = [[] for _ in range(G)]
    for g, h in allowed:                       # successors of each gate, in the set's iteration order
        after[g].append(h)
    triples = [(f, g, h) for f, g in sorted(allowed) for h in after[g]]
    if not use3 or len(triples) > max_triples:
        return bad2, []
    lower |= set(pair_keys)
    return bad2, _cut_windows(triples, [product_key(cs, t) for t in triples], lower)


def cut_list(n: int):
    """bipartitions (A, rest), each once: |A| <= n / 2, and A holds qubit 0 when both halves are equal"""
    out = []
    for r in range(1, n // 2 + 1):
        out.extend(A for A in itertools.combinations(range(n), r) if r != n - r or 0 in A)
    return out


def cx_table(n: int):
    """(cuts, radices, table): the fewest two-qubit gates whose placements cross every cut at least as often as
    that cut's need, for every need vector (need of cut i in [0, 2|A_i|], mixed radix 2|A_i| + 1, cut 0 the
    lowest digit).  A gate lowers the need of each cut it crosses by one, so table[v] = 1 + the least
    table[v'] over the placements that cross a cut with a positive need, v' being v after that placement;
    v' < v in the mixed-radix order, so one ascending pass fills the table.  None (no table) above 2e6 entries."""
    cuts = cut_list(n)
    rad = [2 * len(A) + 1 for A in cuts]
    size = 1
    for r in rad:
        size *= r
    if size > 2_000_000:
        return cuts, rad, None
    stride = [1] * len(cuts)
    for i in range(1, len(cuts)):
        stride[i] = stride[i - 1] * rad[i - 1]
    crossed = [[i for i, A in enumerate(cuts) if (a in A) != (b in A)] for a, b in itertools.combinations(range(n), 2)]
    table = [0] * size
    for v in range(1, size):
        need = [(v // s) % r for s, r in zip(stride, rad)]
        best = None
        for cr in crossed:
            lowered = sum(stride[i] for i in cr if need[i] > 0)
            if lowered and (best is None or table[v - lowered] < best):
                best = table[v - lowered]
        table[v] = 1 + best
    return cuts, rad, table


def _local_entries(g):
    """the gate's local matrix, row-major, as Z[w] entries"""
    _, loc = W._GATES[g]
    L = 1 << arity(g)
    return [W._wpow(loc[(l, m)]) if (l, m) in loc else W.ZERO for l in range(L) for m in range(L)]


def _zs(x) -> str:
    return " ".join(str(v) for v in x)


def spec_text(n: int, circ, depth: int, gates=None, use3: bool = True) -> str:
    """the .spec text of the unitary of circ (spec.py:356-415, unitary targets)"""
    gates = sorted(set(W.CLIFFORD_T) | {g for g, _ in circ}) if gates is None else list(gates)
    cs = candidates(n, gates)
    bad2, bad3 = window_rules(cs, use3)
    k, U = W.unitary(n, circ)
    cuts, _, table = cx_table(n)
    standard = set(gates) <= STANDARD
    lite = not standard and set(gates) <= STANDARD | {"ccx"}
    out = io.StringIO()
    w = out.write
    w("cheese-spec 1\n")
    w(f"n {n}\nmode unitary\nphase exact\ndepth {depth}\n")
    w(f"standard {int(standard)}\n")
    if lite:
        w(f"lite 1\ntdiv {max(PTM_SDE[g] for g in gates if g in PTM_SDE)}\n")
    w(f"gates {len(cs)}\n")
    for name, g, qs in cs:
        ig = DAGGER.get(g, g)
        fwd = " ".join(_zs(x) for x in _local_entries(g))
        inv = " ".join(_zs(x) for x in _local_entries(ig))
        w(f"{name} {len(qs)} {' '.join(str(q) for q in qs)} {W._GATES[g][0]} {fwd} {W._GATES[ig][0]} {inv}\n")
    w(f"bad2 {len(bad2)}\n" + "".join(f"{a} {b}\n" for a, b in bad2))
    w(f"bad3 {len(bad3)}\n" + "".join(f"{a} {b} {c}\n" for a, b, c in bad3))
    w(f"target {k}\n" + "".join(" ".join(_zs(x) for x in row) + "\n" for row in U))
    w(f"cuts {len(cuts)}\n" + "".join(f"{len(A)} {' '.join(str(q) for q in A)}\n" for A in cuts))
    if table is None:
        w("cxtable 0\n")
    else:
        w(f"cxtable {len(table)}\n" + " ".join(str(v) for v in table) + "\n")
    w("end\n")
    return out.getvalue()


def main(argv=None) -> int:
    import argparse
    a = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    a.add_argument("qasm")
    a.add_argument("depth", type=int)
    a.add_argument("--no-w3", action="store_true")
    o = a.parse_args(argv)
    with open(o.qasm) as f:
        n, circ = W.parse_qasm(f.read())
    sys.stdout.write(spec_text(n, circ, o.depth, use3=not o.no_w3))
    return 0


if __name__ == "__main__":
    sys.exit(main())
