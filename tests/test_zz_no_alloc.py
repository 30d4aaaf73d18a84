"""chz_cliff_query_words_noalloc (device buffers, fixed per-device scratch, no allocation) against
chz_cliff_query_words on the same seeded words: two calls in a row, the second smaller, on the table's own stream
and on a caller stream; sizes above the scratch capacity are refused before any device work.  Needs an MI355X."""
import ctypes

import numpy as np
import pytest

from cheese_amd import clifford as CL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def table():
    t = CL.CliffordTable(3)
    yield t
    t.close()


def _noalloc():
    L = CL._lib()
    fn = L.chz_cliff_query_words_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p,
                   ctypes.c_void_p]
    return fn


class _Hip:
    """the HIP runtime the library itself links (device buffers for the test; a second runtime, such as the one a
    framework bundles, would not share the library's device context)"""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]

    def malloc(self, n):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(1, n)) == 0
        return p

    def upload(self, arr):
        p = self.malloc(arr.nbytes)
        assert self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0       # host to device
        return p

    def download(self, p, arr):
        assert self.L.hipMemcpy(arr.ctypes.data, p, arr.nbytes, 2) == 0       # device to host
        return arr


@pytest.mark.parametrize("own_stream", [True, False])
def test_noalloc_query_equals_query_words(table, own_stream):
    hip = _Hip()
    fn = _noalloc()
    stream = None
    if not own_stream:
        stream = ctypes.c_void_p()
        assert hip.L.hipStreamCreate(ctypes.byref(stream)) == 0
    for seed, count in ((11, 20000), (12, 3001)):          # the second call smaller than the first
        words = CL.random_words(seed, count, table.ng)
        want, _ = table.query_words(words)
        flat = np.array([g for w in words for g in w], dtype=np.uint8)
        offs = np.zeros(count + 1, dtype=np.int64)
        offs[1:] = np.cumsum([len(w) for w in words])
        d_letters, d_offs = hip.upload(flat), hip.upload(offs)
        d_dist = hip.upload(np.full(count, -7, dtype=np.int32))
        print(f"noalloc query: {count} words, own stream {own_stream}", flush=True)
        rc = fn(table.h, d_letters, d_offs, count, d_dist, stream)
        assert rc == 0, CL._lib().chz_cliff_last_error().decode()
        if stream is not None:
            assert hip.L.hipStreamSynchronize(stream) == 0
        got = hip.download(d_dist, np.zeros(count, dtype=np.int32)).tolist()
        for p in (d_letters, d_offs, d_dist):
            hip.L.hipFree(p)
        assert got == want
        assert min(got) >= 0
    if stream is not None:
        hip.L.hipStreamDestroy(stream)


def test_noalloc_refuses_more_than_its_scratch(table):
    fn = _noalloc()
    rc = fn(table.h, ctypes.c_void_p(8), ctypes.c_void_p(8), (1 << 20) + 1, ctypes.c_void_p(8), None)
    assert rc == 4
    assert "scratch" in CL._lib().chz_cliff_last_error().decode()
