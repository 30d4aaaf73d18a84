"""Host mirror of the Clifford-group table (the reference's cliffopt table, src/clifford.cu): build the BFS
table on the GPU, batched distance queries of words, optimal-word read-back.  Gate-set letters: for each
qubit q in order, H S Sdg X Z on q; then CX a->b for every ordered pair (a, b), a-major."""
from __future__ import annotations

import ctypes

from . import lib as _lib_loader
from .workloads import SplitMix64

_sig_done = False


def _lib():
    global _sig_done
    L = _lib_loader()
    if not _sig_done:
        P = ctypes.POINTER
        sigs = {
            "chz_cliff_build": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, P(ctypes.c_void_p)]),
            "chz_cliff_levels": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_long), ctypes.c_int]),
            "chz_cliff_num_gates": (ctypes.c_int, [ctypes.c_void_p]),
            "chz_cliff_build_ms": (ctypes.c_double, [ctypes.c_void_p]),
            "chz_cliff_digest": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_uint64), P(ctypes.c_longlong)]),
            "chz_cliff_query_words": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_uint8), P(ctypes.c_longlong),
                                                     ctypes.c_longlong, P(ctypes.c_int), P(ctypes.c_double)]),
            "chz_cliff_word_tableau": (ctypes.c_uint64, [ctypes.c_void_p, P(ctypes.c_uint8), ctypes.c_int]),
            "chz_cliff_readback": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_uint64), ctypes.c_longlong, ctypes.c_int,
                                                  P(ctypes.c_uint8), P(ctypes.c_int)]),
            "chz_cliff_sample": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_longlong, P(ctypes.c_uint64)]),
            "chz_cliff_opt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int,
                                             ctypes.c_char_p, ctypes.c_size_t]),
            "chz_cliff_free": (None, [ctypes.c_void_p]),
            "chz_cliff_last_error": (ctypes.c_char_p, []),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _sig_done = True
    return L


def gate_names(n):
    out = []
    for q in range(n):
        out += [f"h {q}", f"s {q}", f"sdg {q}", f"x {q}", f"z {q}"]
    out += [f"cx {a} {b}" for a in range(n) for b in range(n) if a != b]
    return out


def word_gates(n, word):
    """a gate-set word as a circuit [(gate, qubits)] (letters applied in order)"""
    names = gate_names(n)
    out = []
    for l in word:
        parts = names[l].split()
        out.append((parts[0], tuple(int(q) for q in parts[1:])))
    return out


def random_words(seed, count, ng, lmin=5, lmax=40):
    """seeded random words: length uniform in [lmin, lmax], letters uniform over the ng gates"""
    r = SplitMix64(seed)
    words = []
    for _ in range(count):
        L = lmin + r.below(lmax - lmin + 1)
        words.append([r.below(ng) for _ in range(L)])
    return words


class CliffordTable:
    def __init__(self, n, device=0):
        L = _lib()
        self.n = n
        self.h = ctypes.c_void_p()
        rc = L.chz_cliff_build(n, device, ctypes.byref(self.h))
        if rc:
            raise RuntimeError(f"chz_cliff_build exit {rc}: {L.chz_cliff_last_error().decode()}")
        self.ng = L.chz_cliff_num_gates(self.h)
        self.build_ms = L.chz_cliff_build_ms(self.h)

    def levels(self):
        arr = (ctypes.c_long * 128)()
        k = _lib().chz_cliff_levels(self.h, arr, 128)
        return list(arr[:k])

    def digest(self):
        d, e = ctypes.c_uint64(), ctypes.c_longlong()
        _lib().chz_cliff_digest(self.h, ctypes.byref(d), ctypes.byref(e))
        return d.value, e.value

    def query_words(self, words):
        """distances of many words in one batched launch; returns (list, device ms)"""
        flat = [g for w in words for g in w]
        offs = [0]
        for w in words:
            offs.append(offs[-1] + len(w))
        la = (ctypes.c_uint8 * max(1, len(flat)))(*flat)
        oa = (ctypes.c_longlong * len(offs))(*offs)
        out = (ctypes.c_int * max(1, len(words)))()
        ms = ctypes.c_double()
        rc = _lib().chz_cliff_query_words(self.h, la, oa, len(words), out, ctypes.byref(ms))
        if rc:
            raise RuntimeError(f"query exit {rc}")
        return list(out[:len(words)]), ms.value

    def tableau(self, word):
        arr = (ctypes.c_uint8 * max(1, len(word)))(*word)
        return _lib().chz_cliff_word_tableau(self.h, arr, len(word))

    def readback(self, tableaux, maxlen=32):
        n = len(tableaux)
        ta = (ctypes.c_uint64 * max(1, n))(*tableaux)
        words = (ctypes.c_uint8 * max(1, n * maxlen))()
        lens = (ctypes.c_int * max(1, n))()
        rc = _lib().chz_cliff_readback(self.h, ta, n, maxlen, words, lens)
        if rc:
            raise RuntimeError(f"readback exit {rc}")
        return [list(words[i * maxlen:i * maxlen + lens[i]]) if lens[i] >= 0 else None for i in range(n)]

    def sample(self, seed, count):
        """count uniformly random group elements (tableaux) of the table, seeded"""
        out = (ctypes.c_uint64 * max(count, 1))()
        rc = _lib().chz_cliff_sample(self.h, seed, count, out)
        if rc:
            raise RuntimeError(f"chz_cliff_sample exit {rc}: {_lib().chz_cliff_last_error().decode()}")
        return list(out[:count])

    def optimize(self, job_text: str, beam: int = 16) -> str:
        """cliffopt on a job text (chz_cliff_opt): the reference's stdout, first line JSON, then the gate lines"""
        raw = job_text.encode()
        buf = ctypes.create_string_buffer(1 << 22)
        rc = _lib().chz_cliff_opt(self.h, raw, len(raw), beam, buf, len(buf))
        if rc:
            raise RuntimeError(f"chz_cliff_opt exit {rc}: {_lib().chz_cliff_last_error().decode()}")
        return buf.value.decode()

    def close(self):
        if self.h:
            _lib().chz_cliff_free(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


JOB_GATES = ("h", "s", "sdg", "x", "z", "cx")


def random_job(seed, n=3, t=4, lmin=3, lmax=12):
    """a seeded cliffopt job: t T gates (uniform qubit, uniform dagger) between t + 1 random Clifford segments of
    lmin..lmax gates over h s sdg x z cx (cx on a uniform ordered pair).  Returns (job text, circuit as
    (gate, qubits) for the exact unitary)"""
    r = SplitMix64(seed)
    tq = [(r.below(n), r.below(2)) for _ in range(t)]
    segs = []
    for _ in range(t + 1):
        seg = []
        for _ in range(lmin + r.below(lmax - lmin + 1)):
            g = JOB_GATES[r.below(len(JOB_GATES))]
            if g == "cx":
                a = r.below(n)
                b = r.below(n - 1)
                b += b >= a
                seg.append((g, (a, b)))
            else:
                seg.append((g, (r.below(n),)))
        segs.append(seg)
    lines = [f"{n} {t}"] + [f"{q} {d}" for q, d in tq]
    circ = []
    for j, seg in enumerate(segs):
        lines.append(" ".join([str(len(seg))] + [" ".join([g] + [str(x) for x in qs]) for g, qs in seg]))
        circ += seg
        if j < t:
            circ.append(("tdg" if tq[j][1] else "t", (tq[j][0],)))
    return "\n".join(lines) + "\n", circ


def parse_opt_output(text):
    """the (gate, qubits) circuit and the JSON summary of cliffopt's output"""
    import json
    lines = text.strip().splitlines()
    head = json.loads(lines[0])
    circ = []
    for ln in lines[1:]:
        parts = ln.split()
        circ.append((parts[0], tuple(int(x) for x in parts[1:])))
    return head, circ
