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

    def close(self):
        if self.h:
            _lib().chz_cliff_free(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
