"""Clifford-group table: the oracle pinned by closed forms (CPU), and the HIP table bit-identical to the
oracle's (GPU)."""
import ctypes
import os

import pytest

from conftest import ROOT

from cheese_amd import clifford

ORDERS = {1: 24, 2: 11520, 3: 92897280}     # 2^(n^2+2n) prod_{j=1..n} (4^j - 1), src/clifford.cu:252


@pytest.fixture(scope="session")
def orc():
    path = os.path.join(ROOT, "oracle", "libclifford_oracle.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True, capture_output=True)
    L = ctypes.CDLL(path)
    L.orc_cliff_build.restype = ctypes.c_void_p
    L.orc_cliff_build.argtypes = [ctypes.c_int, ctypes.c_int]
    L.orc_cliff_levels.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long), ctypes.c_int]
    L.orc_cliff_digest.restype = ctypes.c_uint64
    L.orc_cliff_digest.argtypes = [ctypes.c_void_p]
    L.orc_cliff_word.restype = ctypes.c_uint64
    L.orc_cliff_word.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.orc_cliff_find.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    L.orc_cliff_readback.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.orc_cliff_free.argtypes = [ctypes.c_void_p]
    return L


def _levels(L, h):
    arr = (ctypes.c_long * 64)()
    k = L.orc_cliff_levels(h, arr, 64)
    return list(arr[:k])


def _word(L, h, w):
    arr = (ctypes.c_int * max(1, len(w)))(*w)
    return L.orc_cliff_word(h, arr, len(w))


@pytest.mark.parametrize("n", [1, 2])
def test_oracle_group_order_and_identities(orc, n):
    h = orc.orc_cliff_build(n, 0)
    lv = _levels(orc, h)
    assert sum(lv) == ORDERS[n] and lv[0] == 1
    ng = 5 * n + n * (n - 1)
    assert lv[1] == ng                                    # every single gate is a distinct non-identity Clifford
    for q in range(n):
        H, S, Sdg, X, Z = (5 * q + i for i in range(5))
        assert orc.orc_cliff_find(h, _word(orc, h, [H, H])) == 0          # H^2 = I
        assert orc.orc_cliff_find(h, _word(orc, h, [S, Sdg])) == 0        # S Sdg = I
        assert _word(orc, h, [S, S]) == _word(orc, h, [Z])                 # S^2 = Z
        assert _word(orc, h, [H, X, H]) == _word(orc, h, [Z])              # H X H = Z
    # read-back gives a word of the stored length that reproduces the tableau
    import random
    rng = random.Random(1)
    for _ in range(200):
        w = [rng.randrange(ng) for _ in range(rng.randrange(1, 15))]
        t = _word(orc, h, w)
        out = (ctypes.c_int * 32)()
        ln = orc.orc_cliff_readback(h, t, out, 32)
        assert ln == orc.orc_cliff_find(h, t) and ln <= len(w)
        assert _word(orc, h, list(out[:ln])) == t
    orc.orc_cliff_free(h)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2, 3])
def test_hip_table_identical_to_oracle(orc, n):
    print(f"case clifford n={n}", flush=True)
    T = clifford.CliffordTable(n)
    h = orc.orc_cliff_build(n, 0)
    assert T.levels() == _levels(orc, h)
    assert sum(T.levels()) == ORDERS[n]
    dig, entries = T.digest()
    assert entries == ORDERS[n]
    assert dig == orc.orc_cliff_digest(h)
    # batched random queries (words of 5..40 letters) against the oracle's table: at n = 3 the 1e5 queries of the
    # Clifford workload
    words = clifford.random_words(1234 + n, 100000 if n == 3 else 2000, T.ng)
    got, _ = T.query_words(words)
    want = [orc.orc_cliff_find(h, _word(orc, h, w)) for w in words]
    assert got == want
    # optimal words: same rule as the oracle, so the same words
    tabs = [T.tableau(w) for w in words[:500]]
    assert tabs == [_word(orc, h, w) for w in words[:500]]
    back = T.readback(tabs, 32)
    for t, w in zip(tabs, back):
        out = (ctypes.c_int * 32)()
        ln = orc.orc_cliff_readback(h, t, out, 32)
        assert w == list(out[:ln])
    orc.orc_cliff_free(h)
    T.close()
