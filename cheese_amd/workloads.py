"""Seeded synthetic workloads for the gate-count search: random Clifford+T circuits, their exact unitaries
and .spec text built from a per-(n, gate set) skeleton.

Exact arithmetic: an entry a + b w + c w^2 + d w^3 of Z[w] (w = e^{i pi/4}) is a 4-tuple, a matrix is
(k, rows) meaning rows / sqrt2^k, normalised so equal matrices have equal representations (the .spec
convention of python/cheese/spec.py:1-8).  Bit q of a basis index is qubit q.
"""
from __future__ import annotations

import os

MASK64 = (1 << 64) - 1
ZERO = (0, 0, 0, 0)
ONE = (1, 0, 0, 0)

# gate -> (sqrt2 exponent, local matrix as {(row, col): w-power}); local index bit j = qubit qs[j]
_GATES = {
    "x": (0, {(0, 1): 0, (1, 0): 0}),
    "y": (0, {(0, 1): 6, (1, 0): 2}),
    "z": (0, {(0, 0): 0, (1, 1): 4}),
    "s": (0, {(0, 0): 0, (1, 1): 2}),
    "sdg": (0, {(0, 0): 0, (1, 1): 6}),
    "t": (0, {(0, 0): 0, (1, 1): 1}),
    "tdg": (0, {(0, 0): 0, (1, 1): 7}),
    "h": (1, {(0, 0): 0, (0, 1): 0, (1, 0): 0, (1, 1): 4}),
    "cx": (0, {(0, 0): 0, (1, 3): 0, (2, 2): 0, (3, 1): 0}),
    "cz": (0, {(0, 0): 0, (1, 1): 0, (2, 2): 0, (3, 3): 4}),
    "ccx": (0, {(l, {3: 7, 7: 3}.get(l, l)): 0 for l in range(8)}),
}
CLIFFORD_T = ("x", "z", "h", "s", "sdg", "t", "tdg", "cx")


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): the seeded generator of every synthetic workload here"""

    def __init__(self, seed: int):
        self.s = seed & MASK64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        """uniform integer in [0, n) by rejection"""
        lim = (1 << 64) - ((1 << 64) % n)
        while True:
            x = self.next()
            if x < lim:
                return x % n


def random_circuit(seed: int, n: int = 3, lmin: int = 10, lmax: int = 13, gates=CLIFFORD_T):
    """length uniform in [lmin, lmax]; each gate uniform over `gates`; a one-qubit gate on a uniform qubit,
    cx on a uniform ordered pair of distinct qubits"""
    r = SplitMix64(seed)
    L = lmin + r.below(lmax - lmin + 1)
    circ = []
    for _ in range(L):
        g = gates[r.below(len(gates))]
        if g in ("cx", "cz"):
            a = r.below(n)
            b = r.below(n - 1)
            b += b >= a
            circ.append((g, (a, b)))
        else:
            circ.append((g, (r.below(n),)))
    return circ


def _zmul(x, y):
    a, b, c, d = x
    e, f, g, h = y
    return (a * e - b * h - c * g - d * f, a * f + b * e - c * h - d * g,
            a * g + b * f + c * e - d * h, a * h + b * g + c * f + d * e)


def _wpow(p):
    p %= 8
    out = [0, 0, 0, 0]
    out[p % 4] = 1 if p < 4 else -1
    return tuple(out)


def _down(x):
    a, b, c, d = x
    if (a - c) % 2 or (b - d) % 2:
        return None
    return ((b - d) // 2, (a + c) // 2, (b + d) // 2, (c - a) // 2)


def normalise(k, M):
    while k > 0:
        D = [[_down(x) for x in row] for row in M]
        if any(x is None for row in D for x in row):
            break
        M, k = D, k - 1
    return k, [list(r) for r in M]


def apply_gate(U, g, qs, n):
    """(k, M) -> gate * M, normalised"""
    k, M = U
    gk, loc = _GATES[g]
    N = 1 << n
    mask = sum(1 << q for q in qs)
    out = [[ZERO] * N for _ in range(N)]
    for r in range(N):
        lr = sum(((r >> q) & 1) << j for j, q in enumerate(qs))
        base = r & ~mask
        for (l, m), p in loc.items():
            if l != lr:
                continue
            x = base | sum(((m >> j) & 1) << q for j, q in enumerate(qs))
            wp = _wpow(p)
            for c in range(N):
                v = M[x][c]
                if v != ZERO:
                    t = _zmul(wp, v)
                    o = out[r][c]
                    out[r][c] = (o[0] + t[0], o[1] + t[1], o[2] + t[2], o[3] + t[3])
    return normalise(k + gk, out)


def unitary(n, circ):
    N = 1 << n
    U = (0, [[ONE if i == j else ZERO for j in range(N)] for i in range(N)])
    for g, qs in circ:
        U = apply_gate(U, g, qs, n)
    return U


def target_block(U) -> str:
    k, M = U
    return f"target {k}\n" + "".join(" ".join(" ".join(map(str, x)) for x in row) + "\n" for row in M)


SKELETON_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "skeletons")


def skeleton(n: int) -> str:
    with open(os.path.join(SKELETON_DIR, f"n{n}.spec")) as f:
        return f.read()


def spec_from_skeleton(skel: str, U, depth: int) -> str:
    """the skeleton's text with its depth line and target block replaced"""
    lines = skel.splitlines(keepends=True)
    out, i = [], 0
    N = len(U[1])
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("depth "):
            out.append(f"depth {depth}\n")
        elif ln.startswith("target "):
            out.append(target_block(U))
            i += N
        else:
            out.append(ln)
        i += 1
    return "".join(out)


def random_spec(seed: int, n: int = 3, lmin: int = 10, lmax: int = 13):
    """(circuit, spec text at budget = its length) for a seeded random Clifford+T circuit"""
    circ = random_circuit(seed, n, lmin, lmax)
    return circ, spec_from_skeleton(skeleton(n), unitary(n, circ), len(circ))


def parse_qasm(text: str):
    """qubit count and (gate, qubits) list of a small OpenQASM 2/3 file (measurements dropped)"""
    import re
    n, circ = 0, []
    for line in text.splitlines():
        line = line.split("//")[0].strip()
        m = re.match(r"(qreg\s+\w+\[(\d+)\]|qubit\[(\d+)\])", line)
        if m:
            n = int(m.group(2) or m.group(3))
            continue
        m = re.match(r"([a-z]+)\s+(.*);", line)
        if m and m.group(1) in _GATES:
            circ.append((m.group(1), tuple(int(q) for q in re.findall(r"\[(\d+)\]", m.group(2)))))
    return n, circ
