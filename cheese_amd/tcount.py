"""T-count mode, host side: the .tspec text of a target and the exact check of a decomposition (both computed by
the library's host code, chz_tspec_write / chz_tc_verify), seeded T-count workloads, and the GPU coset table.

A unitary U is U = R(P_1) ... R(P_t) C with R(P) = ((1 + w) I + (1 - w) P) / 2, C a Clifford, t the T-count,
global phase ignored (python/cheese/tcount.py:1-12).  The search runs on channel representations
Uc[a][b] = Tr(P_a U P_b U^dag) / 2^n: real 4^n x 4^n matrices with entries (x + y sqrt2) / sqrt2^e.
Matrices over Z[w] are (k, rows) as in workloads.py.
"""
from __future__ import annotations

from . import workloads as W


def _entries(n, U):
    """(k, rows) -> (k, the N x N entries as a flat int64 array of a b c d coefficients)"""
    import ctypes
    k, M = U
    N = 1 << n
    if len(M) != N or any(len(r) != N for r in M):
        raise ValueError(f"expected a {N} x {N} matrix")
    flat = [c for r in M for z in r for c in z]
    return k, (ctypes.c_longlong * len(flat))(*flat)


def tspec_text(n, U):
    """the .tspec text of the unitary U = (k, rows) (chz_tspec_write: transfer matrix and rotation tables; the
    text `python3 -m cheese.tcount spec` writes, python/cheese/tcount.py:157-167)"""
    import ctypes
    L = _tc_lib()
    k, ent = _entries(n, U)
    size = L.chz_tspec_write(n, k, ent, None, 0)
    if size < 0:
        raise ValueError(f"chz_tspec_write: {size}")
    buf = ctypes.create_string_buffer(size + 1)
    got = L.chz_tspec_write(n, k, ent, buf, size + 1)
    if got != size:
        raise ValueError(f"chz_tspec_write: {got}")
    return buf.value.decode()


def verify(n, U, rotations):
    """exact check of a T-count answer (chz_tc_verify): peel the listed rotations (p, inv) off U in order,
    U = X_1 ... X_t C with X = R(P_p) or R(P_p)^-1, and test that a Clifford remains"""
    import ctypes
    L = _tc_lib()
    k, ent = _entries(n, U)
    flat = [int(v) for pr in rotations for v in pr]
    rots = (ctypes.c_int * max(1, len(flat)))(*flat)
    r = L.chz_tc_verify(n, k, ent, rots, len(rotations))
    if r < 0:
        raise ValueError(f"chz_tc_verify: {r}")
    return r == 1


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
            "chz_tspec_write": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P(ctypes.c_longlong), ctypes.c_char_p,
                                               ctypes.c_size_t]),
            "chz_tc_verify": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P(ctypes.c_longlong), P(ctypes.c_int),
                                             ctypes.c_int]),
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
