"""The request server: a seeded mixed stream (gate-count, T-count, Clifford re-synthesis) answered in served mode
(continuous batching, overlapping workers) gives, request by request, the answers of plain mode (one request at a
time, alone on the engines); every answer is checked exactly against its request."""
import pytest

from conftest import read_spec, read_tspec

import cheese_amd
from cheese_amd import clifford, serve
from cheese_amd import tcount as T
from cheese_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream():
    cl = clifford.CliffordTable(3)
    try:
        reqs = serve.mixed_requests(48, seed=3, cliff=cl)
        tabs = {i: cl.tableau(r["word"]) for i, r in enumerate(reqs) if r["kind"] == serve.CLIFFORD}
        dists = {i: d for i, d in zip(tabs, cl.query_words([reqs[i]["word"] for i in tabs])[0])}
        return reqs, tabs, dists
    finally:
        cl.close()


def _run(reqs, plain, arrivals=None, max_batch=0):
    spec2 = W.spec_from_skeleton(W.skeleton(2), W.unitary(2, []), 1)
    spec3 = W.spec_from_skeleton(W.skeleton(3), W.unitary(3, []), 1)
    s = serve.Server(spec2, spec3, read_tspec("toffoli"), plain=plain, max_batch=max_batch)
    try:
        answers, _ = serve.replay(s, reqs, arrivals or [0.0] * len(reqs))
        assert s.error() == ""
    finally:
        s.close()
    return answers


def test_served_answers_equal_plain_answers(stream):
    reqs, tabs, dists = stream
    plain = _run(reqs, plain=True)
    arrivals = serve.poisson_arrivals(len(reqs), rate=400.0, seed=9)
    served = _run(reqs, plain=False, arrivals=arrivals)
    cl = clifford.CliffordTable(3)
    try:
        for i, r in enumerate(reqs):
            a, b = plain[i], served[i]
            for f in ("rc", "len", "out", "found", "split"):
                assert a[f] == b[f], (i, r["kind"], f)
            if r["kind"] == serve.GATES:
                assert b["rc"] == 0 and b["len"] <= r["length"]
                assert cheese_amd.verify_spec(r["text"], b["out"])
            elif r["kind"] == serve.TCOUNT:
                assert b["rc"] == 0 and b["len"] <= r["k"]
                rots = [[b["out"][2 * k], b["out"][2 * k + 1]] for k in range(b["len"])]
                assert T.verify(3, W.unitary(3, r["circ"]), rots)
            else:
                assert b["rc"] == 0 and b["len"] == dists[i]
                assert cl.tableau(b["out"]) == tabs[i]
    finally:
        cl.close()


def test_capped_batches_join_as_slots_free(stream):
    """a burst offered at once to batches of at most 4 live searches (each T-count batch at most 4 targets): the
    queued searches join as earlier ones end, over many steps, and every answer is still the plain answer"""
    reqs, _, _ = stream
    plain = _run(reqs, plain=True)
    served = _run(reqs, plain=False, max_batch=4)
    for i, r in enumerate(reqs):
        for f in ("rc", "len", "out", "found", "split"):
            assert plain[i][f] == served[i][f], (i, r["kind"], f)


def test_two_thousand_request_trace():
    """C5 at its stated size: the seeded 2,000-request mixed stream (981 gate-count, 728 T-count, 291 Clifford) on
    Poisson arrivals; every served answer equals the plain answer and is checked exactly against its request, no
    device error, and the measured device memory stays under the HBM size"""
    cl = clifford.CliffordTable(3)
    try:
        reqs = serve.mixed_requests(2000, seed=1, cliff=cl)
        tabs = {i: cl.tableau(r["word"]) for i, r in enumerate(reqs) if r["kind"] == serve.CLIFFORD}
        dists = {i: d for i, d in zip(tabs, cl.query_words([reqs[i]["word"] for i in tabs])[0])}
    finally:
        cl.close()
    print("2000-request trace: plain", flush=True)
    plain = _run(reqs, plain=True)
    arrivals = serve.poisson_arrivals(len(reqs), rate=375.0, seed=1)
    print("2000-request trace: served", flush=True)
    spec2 = W.spec_from_skeleton(W.skeleton(2), W.unitary(2, []), 1)
    spec3 = W.spec_from_skeleton(W.skeleton(3), W.unitary(3, []), 1)
    s = serve.Server(spec2, spec3, read_tspec("toffoli"), plain=False)
    try:
        served, _ = serve.replay(s, reqs, arrivals)
        assert s.error() == ""
        assert 0 < s.peak_device_bytes < 288 * 2**30
    finally:
        s.close()
    cl = clifford.CliffordTable(3)
    try:
        for i, r in enumerate(reqs):
            a, b = plain[i], served[i]
            for f in ("rc", "len", "out", "found", "split"):
                assert a[f] == b[f], (i, r["kind"], f)
            if r["kind"] == serve.GATES:
                assert b["rc"] == 0 and b["len"] <= r["length"]
                assert cheese_amd.verify_spec(r["text"], b["out"])
            elif r["kind"] == serve.TCOUNT:
                assert b["rc"] == 0 and b["len"] <= r["k"]
                rots = [[b["out"][2 * k], b["out"][2 * k + 1]] for k in range(b["len"])]
                assert T.verify(3, W.unitary(3, r["circ"]), rots)
            else:
                assert b["rc"] == 0 and b["len"] == dists[i]
                assert cl.tableau(b["out"]) == tabs[i]
    finally:
        cl.close()
