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


def _idle(M, bit):
    """the matrix acts as the identity on local qubit `bit`"""
    N = len(M)
    for r in range(N):
        for c in range(N):
            if (r ^ c) & bit:
                if M[r][c] != W.ZERO:
                    return False
            elif not (r | c) & bit and M[r][c] != M[r | bit][c | bit]:
                return False
    return True


def product_key(cs, seq):
    """exact key of the product of candidates seq (first applied first) on the qubits it acts on nontrivially:
    (qubits, sqrt2 exponent, normalised entries) (spec.py:139-163)"""
    qubits = sorted({q for g in seq for q in cs[g][2]})
    k, M = W.unitary(len(qubits), [(cs[g][1], tuple(qubits.index(q) for q in cs[g][2])) for g in seq])
    for i in range(len(qubits) - 1, -1, -1):       # drop qubits the product leaves alone, highest first
        bit = 1 << i
        if _idle(M, bit):
            keep = [r for r in range(len(M)) if not r & bit]
            k, M = W.normalise(k, [[M[r][c] for c in keep] for r in keep])
            del qubits[i]
    return tuple(qubits), k, tuple(tuple(row) for row in M)


def window_rules(cs, use3: bool = True, max_triples: int = 300_000):
    """W2 pairs and W3 triples (spec.py:165-185): a window is cut when its product equals a shorter product, or
    an earlier window of the same length (pairs in lexicographic order; triples over the allowed pairs in sorted
    order, each pair's successors in the iteration order of the allowed-pair set, as the reference enumerates
    them: which member of an equal product survives depends on it)"""
    G = len(cs)
    shorter = {product_key(cs, ())} | {product_key(cs, (g,)) for g in range(G)}
    seen, bad2, pair_keys = {}, [], set()
    for pair in itertools.product(range(G), repeat=2):
        k = product_key(cs, pair)
        pair_keys.add(k)
        if k in shorter or seen.setdefault(k, pair) != pair:
            bad2.append(pair)
    allowed = set(itertools.product(range(G), repeat=2)) - set(bad2)
    order = list(allowed)
    succ = {g: [h for g2, h in order if g2 == g] for g in range(G)}
    triples = [(f, g, h) for f, g in sorted(allowed) for h in succ[g]]
    bad3 = []
    if use3 and len(triples) <= max_triples:
        shorter |= pair_keys
        seen = {}
        for tri in triples:
            k = product_key(cs, tri)
            if k in shorter or seen.setdefault(k, tri) != tri:
                bad3.append(tri)
    return bad2, bad3


def cut_list(n: int):
    """bipartitions (A, rest), each once: |A| <= n / 2, and A holds qubit 0 when both halves are equal"""
    out = []
    for r in range(1, n // 2 + 1):
        out.extend(A for A in itertools.combinations(range(n), r) if r != n - r or 0 in A)
    return out


def cx_table(n: int):
    """(cuts, radices, table): table[needs in mixed radix 2|A|+1] = fewest two-qubit gates whose crossings meet
    every cut's need, by enumeration of per-pair counts and a superset minimum (spec.py:199-226); None when the
    table would exceed 2e6 entries"""
    cuts = cut_list(n)
    rad = [2 * len(A) + 1 for A in cuts]
    size = 1
    for r in rad:
        size *= r
    if size > 2_000_000:
        return cuts, rad, None
    pairs = list(itertools.combinations(range(n), 2))
    top = max(rad) - 1
    best = [1 << 30] * size
    for counts in itertools.product(range(top + 1), repeat=len(pairs)):
        idx, mult = 0, 1
        for A, r in zip(cuts, rad):
            cross = sum(c for (a, b), c in zip(pairs, counts) if (a in A) != (b in A))
            idx += min(cross, r - 1) * mult
            mult *= r
        best[idx] = min(best[idx], sum(counts))
    mult = 1
    for r in rad:
        for i in range(size - 1, -1, -1):
            if (i // mult) % r < r - 1:
                best[i] = min(best[i], best[i + mult])
        mult *= r
    return cuts, rad, best


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
