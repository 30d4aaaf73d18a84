"""T-count mode, host side: channel representations, Pauli rotation tables, .tspec text, exact verification of
decompositions, and seeded T-count workloads.

A unitary U is U = R(P_1) ... R(P_t) C with R(P) = ((1 + w) I + (1 - w) P) / 2, C a Clifford, t the T-count,
global phase ignored (python/cheese/tcount.py:1-12).  The search runs on channel representations
Uc[a][b] = Tr(P_a U P_b U^dag) / 2^n: real 4^n x 4^n matrices with entries (x + y sqrt2) / sqrt2^e.
Exact arithmetic throughout; matrices over Z[w] are (k, rows) as in workloads.py.
"""
from __future__ import annotations

from . import workloads as W

ZERO = W.ZERO
ONE = W.ONE


def _add(x, y):
    return (x[0] + y[0], x[1] + y[1], x[2] + y[2], x[3] + y[3])


def _conj(x):
    return (x[0], -x[3], -x[2], -x[1])


def matmul(A, B):
    (ka, a), (kb, b) = A, B
    N = len(a)
    out = [[ZERO] * N for _ in range(N)]
    for i in range(N):
        for l in range(N):
            if a[i][l] != ZERO:
                ail = a[i][l]
                row = b[l]
                o = out[i]
                for j in range(N):
                    if row[j] != ZERO:
                        o[j] = _add(o[j], W._zmul(ail, row[j]))
    return W.normalise(ka + kb, out)


def dagger(A):
    k, a = A
    N = len(a)
    return k, [[_conj(a[j][i]) for j in range(N)] for i in range(N)]


def pauli_matrix(n, p):
    """P_p, digits per qubit (0 I, 1 X, 2 Y, 3 Z): P|x> = w^ph |x ^ flip>"""
    N = 1 << n
    M = [[ZERO] * N for _ in range(N)]
    for x in range(N):
        flip = ph = 0
        for q in range(n):
            d, bit = (p >> (2 * q)) & 3, (x >> q) & 1
            if d == 1:
                flip |= 1 << q
            elif d == 2:
                flip |= 1 << q
                ph += 2 + 4 * bit
            elif d == 3:
                ph += 4 * bit
        M[x ^ flip][x] = W._wpow(ph)
    return 0, M


def channel(n, U):
    """(e, rows of (x, y)): Uc[a][b] = (x + y sqrt2) / sqrt2^e, normalised"""
    P = 4 ** n
    paulis = [pauli_matrix(n, p) for p in range(P)]
    Ud = dagger(U)
    rows = [[None] * P for _ in range(P)]
    for b in range(P):
        A = matmul(matmul(U, paulis[b]), Ud)
        for a in range(P):
            k, M = matmul(paulis[a], A)
            tr = ZERO
            for i in range(len(M)):
                tr = _add(tr, M[i][i])
            aa, bb, cc, dd = tr
            assert cc == 0 and dd == -bb, "trace not real"
            rows[a][b] = (k, (aa, bb))
    E = max(k for r in rows for k, _ in r) + 2 * n
    out = []
    for r in rows:
        row = []
        for k, (x, y) in r:
            for _ in range(E - 2 * n - k):        # times sqrt2: (x + y s) s = 2y + x s
                x, y = 2 * y, x
            row.append((x, y))
        out.append(row)
    e = E
    while e > 0 and all(x % 2 == 0 for r in out for x, _ in r):   # (x + y s) / s = y + (x / 2) s
        out = [[(y, x // 2) for x, y in r] for r in out]
        e -= 1
    return e, out


def pauli_tables(n):
    """for rotation P_p (p >= 1) and Pauli P_c: (-1, 0) if they commute, else (r, s) with i P_p P_c = s P_r"""
    P = 4 ** n
    paulis = [pauli_matrix(n, p) for p in range(P)]
    iw = (0, 0, 1, 0)
    tab = []
    for p in range(1, P):
        row = []
        for c in range(P):
            _, M = matmul(paulis[p], paulis[c])
            _, M2 = matmul(paulis[c], paulis[p])
            if M == M2:
                row.append((-1, 0))
                continue
            M = [[W._zmul(iw, x) for x in r] for r in M]
            hit = None
            for r in range(P):
                Q = paulis[r][1]
                for sign in (1, -1):
                    if all(M[i][j] == tuple(sign * v for v in Q[i][j]) for i in range(len(M)) for j in range(len(M))):
                        hit = (r, sign)
                if hit:
                    break
            assert hit, "i P Q is not a signed Pauli"
            row.append(hit)
        tab.append(row)
    return tab


_TABLES = {}


def tspec_text(n, U):
    """the .tspec text of U (format of python/cheese/tcount.py:157-167)"""
    e, M = channel(n, U)
    if n not in _TABLES:
        _TABLES[n] = pauli_tables(n)
    tab = _TABLES[n]
    lines = [f"tcount-spec 1\nn {n}\nchannel {e}\n"]
    lines += [" ".join(f"{x} {y}" for x, y in r) + "\n" for r in M]
    lines.append(f"rotations {len(tab)}\n")
    lines += [" ".join(f"{r} {s}" for r, s in row) + "\n" for row in tab]
    lines.append("end\n")
    return "".join(lines)


def rot(n, p):
    """R(P_p) = ((1 + w) I + (1 - w) P_p) / 2"""
    N = 1 << n
    _, Pm = pauli_matrix(n, p)
    a, b = (1, 1, 0, 0), (1, -1, 0, 0)
    return W.normalise(2, [[_add(a if i == j else ZERO, W._zmul(b, Pm[i][j])) for j in range(N)] for i in range(N)])


def _phase_match(A, Q):
    return any(all(A[i][j] == W._zmul(W._wpow(ph), Q[i][j]) for i in range(len(A)) for j in range(len(A)))
               for ph in range(8))


def is_clifford(n, C):
    """C P C^dag is a Pauli up to phase for every non-identity Pauli P"""
    P = 4 ** n
    paulis = [pauli_matrix(n, p) for p in range(P)]
    Cd = dagger(C)
    for p in range(1, P):
        k, A = matmul(matmul(C, paulis[p]), Cd)
        if k != 0 or not any(_phase_match(A, paulis[r][1]) for r in range(1, P)):
            return False
    return True


def verify(n, U, rotations):
    """peel the listed rotations off U (U = X_1 ... X_t C, X = R(P_p) or R(P_p)^-1 for [p, inv]) and check
    that a Clifford remains (python/cheese/tcount.py:174-183 restated)"""
    Wm = U
    for p, inv in rotations:
        X = rot(n, p)
        Wm = matmul(X if inv else dagger(X), Wm)
    return is_clifford(n, Wm)


def random_target(seed, n=3, kmin=3, kmax=7, clifford_words=None):
    """a seeded T-count workload: U = C_k X_k ... C_1 X_1 C_0 with k uniform in [kmin, kmax], each X_i a T or Tdg
    (uniform) on a uniform qubit and each C_i a Clifford given as a gate word; T-count(U) <= k.
    clifford_words(r) returns the next Clifford as a list of (gate, qubits)."""
    r = W.SplitMix64(seed)
    k = kmin + r.below(kmax - kmin + 1)
    circ = []
    circ += clifford_words(r)
    for _ in range(k):
        circ.append(("t" if r.below(2) == 0 else "tdg", (r.below(n),)))
        circ += clifford_words(r)
    return k, circ


def random_clifford_word(r, n=3, lmin=10, lmax=30):
    """a random Clifford as a word over {h, s, sdg, x, z, cx}: length uniform in [lmin, lmax]"""
    gates = ["h", "s", "sdg", "x", "z", "cx"]
    L = lmin + r.below(lmax - lmin + 1)
    out = []
    for _ in range(L):
        g = gates[r.below(len(gates))]
        if g == "cx":
            a = r.below(n)
            b = r.below(n - 1)
            b += b >= a
            out.append(("cx", (a, b)))
        else:
            out.append((g, (r.below(n),)))
    return out


# ---------------------------------------------------------------------------------------------------------------
# the HIP T-count path (include/cheese_hip.h, chz_tc_*): coset table in HBM, batched searches

TMAX = 32


class TcLevel(__import__("ctypes").Structure):
    import ctypes as _c
    _fields_ = [("t", _c.c_int), ("parents", _c.c_long), ("children", _c.c_long), ("cosets", _c.c_long),
                ("fresh", _c.c_long), ("seconds", _c.c_double)]


class TcResult(__import__("ctypes").Structure):
    import ctypes as _c
    _fields_ = [("rc", _c.c_int), ("tcount", _c.c_int), ("split", _c.c_int * 2), ("rotations", _c.c_int * (2 * TMAX)),
                ("nlevels", _c.c_int), ("levels", TcLevel * (TMAX + 1)), ("seconds", _c.c_double)]


_tc_sig_done = False


def _tc_lib():
    import ctypes
    from . import KStat, lib as _lib_loader
    global _tc_sig_done
    L = _lib_loader()
    if not _tc_sig_done:
        P = ctypes.POINTER
        sigs = {
            "chz_tc_table": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                            P(ctypes.c_void_p)]),
            "chz_tc_table_levels": (ctypes.c_int, [ctypes.c_void_p, P(TcLevel), ctypes.c_int]),
            "chz_tc_table_size": (ctypes.c_longlong, [ctypes.c_void_p]),
            "chz_tc_table_depth": (ctypes.c_int, [ctypes.c_void_p]),
            "chz_tc_num_qubits": (ctypes.c_int, [ctypes.c_void_p]),
            "chz_tc_search": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_char_p), P(ctypes.c_size_t), ctypes.c_int,
                                             ctypes.c_int, P(TcResult)]),
            "chz_tc_kernel_stats": (ctypes.c_int, [ctypes.c_void_p, P(KStat), ctypes.c_int]),
            "chz_tc_dedupe_counts": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_double)]),
            "chz_tc_free": (None, [ctypes.c_void_p]),
            "chz_tc_last_error": (ctypes.c_char_p, []),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _tc_sig_done = True
    return L


def level_line(direction, lv):
    """one level line of the reference (src/tcount.cu:385-390)"""
    return ('{"dir": "%s", "t": %d, "parents": %d, "children": %d, "cosets": %d, "new": %d, "seconds": %.2f}'
            % (direction, lv["t"], lv["parents"], lv["children"], lv["cosets"], lv["new"], lv["seconds"]))


def result_line(r):
    """the final line of a search (src/tcount.cu:521-530)"""
    if r["rc"] == 0:
        rots = ", ".join("[%d, %d]" % (p, i) for p, i in r["rotations"])
        return ('{"tcount": %d, "rotations": [%s], "split": [%d, %d], "seconds": %.2f}'
                % (r["tcount"], rots, r["split"][0], r["split"][1], r["seconds"]))
    return '{"tcount_above": %d, "seconds": %.2f}' % (r["maxt"], r["seconds"])


def _level_dict(lv):
    return {"t": lv.t, "parents": lv.parents, "children": lv.children, "cosets": lv.cosets, "new": lv.fresh,
            "seconds": lv.seconds}


class TcError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"rc {rc}: {msg}")
        self.rc = rc


class CosetTable:
    """the table of right cosets with T-count <= K (built on the GPU from any .tspec of the wanted qubit
    count: only its n and rotation tables are used) and batched searches against it"""

    def __init__(self, tspec_text, K=4, device=0):
        import ctypes
        L = _tc_lib()
        b = tspec_text.encode() if isinstance(tspec_text, str) else tspec_text
        h = ctypes.c_void_p()
        rc = L.chz_tc_table(b, len(b), K, device, ctypes.byref(h))
        if rc:
            raise TcError(rc, L.chz_tc_last_error().decode())
        self._h = h
        self.K = K

    @property
    def levels(self):
        L = _tc_lib()
        buf = (TcLevel * (TMAX + 1))()
        n = L.chz_tc_table_levels(self._h, buf, TMAX + 1)
        return [_level_dict(buf[i]) for i in range(n)]

    @property
    def size(self):
        return _tc_lib().chz_tc_table_size(self._h)

    @property
    def depth(self):
        return _tc_lib().chz_tc_table_depth(self._h)

    def table_lines(self):
        return [level_line("table", lv) for lv in self.levels]

    def search(self, tspecs, maxt=16):
        """results of a batch of targets: dicts with rc, tcount, rotations [(p, inv)], split, levels, lines"""
        import ctypes
        L = _tc_lib()
        bs = [t.encode() if isinstance(t, str) else t for t in tspecs]
        n = len(bs)
        arr = (ctypes.c_char_p * max(n, 1))(*bs)
        lens = (ctypes.c_size_t * max(n, 1))(*[len(b) for b in bs])
        res = (TcResult * max(n, 1))()
        rc = L.chz_tc_search(self._h, arr, lens, n, maxt, res)
        if rc:
            raise TcError(rc, L.chz_tc_last_error().decode())
        out = []
        for i in range(n):
            r = res[i]
            d = {"rc": r.rc, "tcount": r.tcount, "split": [r.split[0], r.split[1]], "maxt": maxt,
                 "rotations": [(r.rotations[2 * k], r.rotations[2 * k + 1]) for k in range(r.tcount)] if r.rc == 0 else [],
                 "levels": [_level_dict(r.levels[k]) for k in range(r.nlevels)], "seconds": r.seconds}
            d["lines"] = [level_line("search", lv) for lv in d["levels"]] + ([result_line(d)] if r.rc in (0, 1) else [])
            out.append(d)
        return out

    def kernel_stats(self):
        from . import KStat
        ks = (KStat * 16)()
        n = _tc_lib().chz_tc_kernel_stats(self._h, ks, 16)
        return {ks[i].name.decode(): {"launches": ks[i].launches, "ms": ks[i].ms, "work": ks[i].work} for i in range(n)}

    def dedupe_counts(self):
        import ctypes
        out = (ctypes.c_double * 4)()
        _tc_lib().chz_tc_dedupe_counts(self._h, out)
        return dict(zip(("runs", "visited", "compared", "duplicates"), [int(x) for x in out]))

    def close(self):
        if getattr(self, "_h", None):
            _tc_lib().chz_tc_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def c3_targets(cliff, count=256, seed0=0, n=3, kmin=3, kmax=7):
    """seeded batched T-count workload: target s is U = C_k X_k ... C_1 X_1 C_0 with k uniform in [kmin, kmax],
    each X_i a T or Tdg (uniform) on a uniform qubit and each C_i a uniformly random n-qubit Clifford drawn from
    the Clifford table `cliff` and realised by its optimal word; the T-count is at most k.
    Returns [(k, circuit)]."""
    from . import clifford as CL
    out = []
    for s in range(count):
        r = W.SplitMix64(seed0 + s)
        k = kmin + r.below(kmax - kmin + 1)
        tabs = cliff.sample(((seed0 + s) * 0x9E3779B97F4A7C15 + 1) & ((1 << 64) - 1), k + 1)
        words = cliff.readback(tabs, maxlen=48)
        circ = []
        for i in range(k + 1):
            circ += CL.word_gates(n, words[i])
            if i < k:
                circ.append(("t" if r.below(2) == 0 else "tdg", (r.below(n),)))
        out.append((k, circ))
    return out
