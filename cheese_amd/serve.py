"""Host mirror of the request server (chz_server_*): the mixed request stream, its seeded Poisson trace, and a
driver that submits each request at its arrival time and collects answers.  Served (continuous batching) and
plain (one at a time) modes run the same trace."""
from __future__ import annotations

import ctypes
import math
import time

from . import lib as _lib_loader
from . import workloads as W

GATES, TCOUNT, CLIFFORD = 0, 1, 2
DMAX = 48


class Opts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("plain", ctypes.c_int), ("tcount_K", ctypes.c_int),
                ("max_batch", ctypes.c_int), ("mem_budget_gb", ctypes.c_double)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("id", ctypes.c_uint64), ("text", ctypes.c_char_p), ("len", ctypes.c_size_t),
                ("budget", ctypes.c_int), ("word", ctypes.POINTER(ctypes.c_uint8)), ("wlen", ctypes.c_int)]


class Response(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("kind", ctypes.c_int), ("rc", ctypes.c_int), ("len", ctypes.c_int),
                ("out", ctypes.c_int * (2 * DMAX)), ("found", ctypes.c_long), ("split", ctypes.c_int * 2),
                ("t_submit", ctypes.c_double), ("t_done", ctypes.c_double)]


_done = False


def _lib():
    global _done
    L = _lib_loader()
    if not _done:
        P = ctypes.POINTER
        sigs = {
            "chz_server_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t,
                                               ctypes.c_char_p, ctypes.c_size_t, P(Opts), P(ctypes.c_void_p)]),
            "chz_server_submit": (ctypes.c_int, [ctypes.c_void_p, P(Request)]),
            "chz_server_poll": (ctypes.c_int, [ctypes.c_void_p, P(Response), ctypes.c_int]),
            "chz_server_outstanding": (ctypes.c_longlong, [ctypes.c_void_p]),
            "chz_server_peak_bytes": (ctypes.c_longlong, [ctypes.c_void_p]),
            "chz_server_peak_device_bytes": (ctypes.c_longlong, [ctypes.c_void_p]),
            "chz_server_stats": (ctypes.c_int, [ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int]),
            "chz_server_error": (ctypes.c_char_p, [ctypes.c_void_p]),
            "chz_server_clock": (ctypes.c_double, [ctypes.c_void_p]),
            "chz_server_close": (None, [ctypes.c_void_p]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _done = True
    return L


class Server:
    def __init__(self, spec2: str, spec3: str, tspec3: str, plain=False, device=0, tcount_K=4, max_batch=0,
                 mem_budget_gb=200.0):
        L = _lib()
        o = Opts(device, int(plain), tcount_K, max_batch, mem_budget_gb)
        self._h = ctypes.c_void_p()
        a, b, c = spec2.encode(), spec3.encode(), tspec3.encode()
        rc = L.chz_server_open(a, len(a), b, len(b), c, len(c), ctypes.byref(o), ctypes.byref(self._h))
        if rc:
            raise RuntimeError(f"chz_server_open exit {rc}")

    def submit(self, rid: int, req: dict):
        L = _lib()
        r = Request()
        r.kind = req["kind"]
        r.id = rid
        r.budget = req.get("budget", -1)
        keep = []
        if "text" in req:
            raw = req["text"].encode()
            keep.append(raw)
            r.text = raw
            r.len = len(raw)
        if "word" in req:
            arr = (ctypes.c_uint8 * max(1, len(req["word"])))(*req["word"])
            keep.append(arr)
            r.word = arr
            r.wlen = len(req["word"])
        return L.chz_server_submit(self._h, ctypes.byref(r))

    def poll(self):
        L = _lib()
        buf = (Response * 256)()
        out = []
        while True:
            n = L.chz_server_poll(self._h, buf, 256)
            for i in range(n):
                x = buf[i]
                out.append({"id": x.id, "kind": x.kind, "rc": x.rc, "out": list(x.out[:max(0, x.len) * (2 if x.kind == TCOUNT else 1)]),
                            "len": x.len, "found": x.found, "split": [x.split[0], x.split[1]],
                            "t_submit": x.t_submit, "t_done": x.t_done})
            if n < 256:
                return out

    @property
    def outstanding(self):
        return _lib().chz_server_outstanding(self._h)

    @property
    def peak_bytes(self):
        return _lib().chz_server_peak_bytes(self._h)

    @property
    def peak_device_bytes(self):
        """largest device memory in use seen after a gate-count step (measured)"""
        return _lib().chz_server_peak_device_bytes(self._h)

    def stats(self):
        """per-engine work so far: level steps, seconds and kernel milliseconds of the 2- and 3-qubit gate batches,
        T-count and Clifford batches and seconds"""
        buf = (ctypes.c_double * 16)()
        n = _lib().chz_server_stats(self._h, buf, 16)
        v = list(buf[:n])
        return {"gates2": {"steps": int(v[0]), "seconds": round(v[2], 4), "kernel_ms": round(v[4], 2),
                           "live_peak": int(v[6])},
                "gates3": {"steps": int(v[1]), "seconds": round(v[3], 4), "kernel_ms": round(v[5], 2),
                           "live_peak": int(v[7])},
                "tcount": {"batches": int(v[8]), "seconds": round(v[9], 4)},
                "clifford": {"batches": int(v[10]), "seconds": round(v[11], 4)}}

    def error(self) -> str:
        """the last device error a worker met ("" if none)"""
        return _lib().chz_server_error(self._h).decode()

    def clock(self):
        return _lib().chz_server_clock(self._h)

    def close(self):
        if getattr(self, "_h", None):
            _lib().chz_server_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mixed_requests(n, seed, cliff, ng_cliff=21):
    """the mixed stream: 50% gate-count synthesis of seeded random 2- and 3-qubit Clifford+T circuits of 4..13
    gates at budget = length, 35% T-count (T-count <= k, k in 2..7, Cliffords sampled from the Clifford table),
    15% Clifford re-synthesis of random words of 5..40 letters.  Returns [dict(kind, ...)]"""
    from . import tcount as T
    r = W.SplitMix64(seed)
    kinds = []
    for _ in range(n):
        x = r.below(100)
        kinds.append(GATES if x < 50 else TCOUNT if x < 85 else CLIFFORD)
    ntc = kinds.count(TCOUNT)
    tcs = T.c3_targets(cliff, ntc, seed0=seed * 1000003 + 17, kmin=2, kmax=7)
    out, ti = [], 0
    for i, k in enumerate(kinds):
        s = (seed << 20) + i
        if k == GATES:
            nq = 2 + r.below(2)
            circ, spec = W.random_spec(s, n=nq, lmin=4, lmax=13)
            out.append({"kind": GATES, "text": spec, "budget": -1, "n": nq, "length": len(circ)})
        elif k == TCOUNT:
            kk, circ = tcs[ti]
            ti += 1
            out.append({"kind": TCOUNT, "text": T.tspec_text(3, W.unitary(3, circ)), "budget": 16, "k": kk,
                        "circ": circ})
        else:
            L = 5 + r.below(36)
            out.append({"kind": CLIFFORD, "word": [r.below(ng_cliff) for _ in range(L)]})
    return out


def poisson_arrivals(n, rate, seed):
    r = W.SplitMix64(seed ^ 0x5EED)
    t, out = 0.0, []
    for _ in range(n):
        u = (r.next() >> 11) / float(1 << 53)
        t += -math.log(1.0 - u) / rate
        out.append(t)
    return out


def replay(server: Server, requests, arrivals, deadline_s: float = 600.0):
    """submit request i at arrivals[i] seconds after the start; returns (answers by id, wall seconds).  Raises
    when answers are still missing deadline_s seconds after the last arrival, naming them."""
    t0 = time.perf_counter()
    answers = {}
    i = 0
    n = len(requests)
    base = server.clock()
    while len(answers) < n:
        now = time.perf_counter() - t0
        while i < n and arrivals[i] <= now:
            server.submit(i, requests[i])
            i += 1
        for a in server.poll():
            a["arrival"] = arrivals[a["id"]]
            a["latency"] = a["t_done"] - base - arrivals[a["id"]]
            answers[a["id"]] = a
        if i < n:
            time.sleep(max(0.0, min(0.002, arrivals[i] - (time.perf_counter() - t0))))
        else:
            if time.perf_counter() - t0 > arrivals[-1] + deadline_s:
                missing = [j for j in range(n) if j not in answers]
                raise RuntimeError(f"{len(missing)} requests unanswered {deadline_s:.0f} s after the last arrival "
                                   f"(ids {missing[:20]}); server error: {server.error()!r}")
            time.sleep(0.001)
    wall = time.perf_counter() - t0
    return answers, wall


def service_breakdown(requests, answers):
    """plain mode with every request offered at once serves them in id order: a request's service time is the
    time between its answer and the previous one.  Returns per kind: count, summed service seconds, the largest"""
    ids = sorted(answers)
    prev = min(a["t_submit"] for a in answers.values())
    out = {}
    for i in ids:
        a = answers[i]
        st = a["t_done"] - prev
        prev = a["t_done"]
        k = {GATES: "gates", TCOUNT: "tcount", CLIFFORD: "clifford"}[requests[i]["kind"]]
        d = out.setdefault(k, {"count": 0, "service_s": 0.0, "max_s": 0.0})
        d["count"] += 1
        d["service_s"] += st
        d["max_s"] = max(d["max_s"], st)
    for d in out.values():
        d["service_s"] = round(d["service_s"], 3)
        d["max_s"] = round(d["max_s"], 3)
    return out


def finish_by_kind(requests, answers):
    """per kind, the time from the first submission to the kind's last answer"""
    t0 = min(a["t_submit"] for a in answers.values())
    out = {}
    for i, a in answers.items():
        k = {GATES: "gates", TCOUNT: "tcount", CLIFFORD: "clifford"}[requests[i]["kind"]]
        out[k] = round(max(out.get(k, 0.0), a["t_done"] - t0), 3)
    return out


def summarise(answers, wall):
    lat = sorted(a["latency"] for a in answers.values())

    def pct(q):
        return lat[min(len(lat) - 1, int(q * len(lat)))]
    return {"requests": len(lat), "solves_per_s": round(len(lat) / wall, 2), "wall_s": round(wall, 3),
            "p50_ms": round(1e3 * pct(0.50), 2), "p99_ms": round(1e3 * pct(0.99), 2),
            "mean_ms": round(1e3 * sum(lat) / len(lat), 2)}
