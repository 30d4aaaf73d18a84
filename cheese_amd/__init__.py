"""cheese_amd: MI355X-native exact branch-and-bound synthesis (gate count, T-count tables, Clifford table).

The product is the HIP library lib/libcheese_hip.so behind the C ABI declared in include/cheese_hip.h; this
package is its host-side mirror of the reference's process interface (python/cheese + the `cheese` binary):
the same .spec input, the same per-level JSON lines and the same exit codes.  Loading fails loudly when the
library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CHEESE_AMD_LIB") or os.path.join(HERE, "lib", "libcheese_hip.so")
CLI_PATH = os.path.join(HERE, "bin", "cheese")
DMAX = 48

_lib = None


class Level(ctypes.Structure):
    _fields_ = [("t", ctypes.c_int), ("parents", ctypes.c_long), ("prefixes", ctypes.c_long),
                ("cut_w3", ctypes.c_long), ("cut_lb", ctypes.c_long), ("survivors", ctypes.c_long),
                ("distinct", ctypes.c_long), ("proved_unsat", ctypes.c_int), ("seconds", ctypes.c_double),
                ("dir", ctypes.c_int)]


class Opts(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_int), ("dedupe", ctypes.c_int), ("sym", ctypes.c_int), ("chunk", ctypes.c_long),
                ("profile", ctypes.c_int), ("device", ctypes.c_int), ("table", ctypes.c_int), ("mitm", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("rc", ctypes.c_int), ("len", ctypes.c_int), ("found", ctypes.c_long),
                ("circuit", ctypes.c_int * DMAX), ("seconds", ctypes.c_double), ("split_kind", ctypes.c_int),
                ("split", ctypes.c_int * 2)]


DIRS = {1: "table", 2: "fwd", 3: "bwd", 4: "join"}


class KStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("launches", ctypes.c_long), ("ms", ctypes.c_double),
                ("work", ctypes.c_double)]


class Event(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int), ("kind", ctypes.c_int), ("level", Level), ("result", Result)]


LEVEL_CB = ctypes.CFUNCTYPE(None, ctypes.POINTER(Level), ctypes.c_void_p)

# entry points declared in include/cheese_hip.h, with their ctypes signatures
SIGNATURES = {
    "chz_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(Opts), ctypes.POINTER(ctypes.c_void_p)]),
    "chz_run": (ctypes.c_int, [ctypes.c_void_p, LEVEL_CB, ctypes.c_void_p, ctypes.POINTER(Result)]),
    "chz_num_gates": (ctypes.c_int, [ctypes.c_void_p]),
    "chz_num_qubits": (ctypes.c_int, [ctypes.c_void_p]),
    "chz_depth": (ctypes.c_int, [ctypes.c_void_p]),
    "chz_sym_size": (ctypes.c_int, [ctypes.c_void_p]),
    "chz_gate_name": (ctypes.c_char_p, [ctypes.c_void_p, ctypes.c_int]),
    "chz_close": (None, [ctypes.c_void_p]),
    "chz_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "chz_kernel_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(KStat), ctypes.c_int]),
    "chz_expand_stages": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "chz_time_kernels": (ctypes.c_uint, [ctypes.c_void_p, ctypes.c_uint]),
    "chz_verify": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
    "chz_last_error": (ctypes.c_char_p, []),
    "chz_verify_spec": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
    "chz_tune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    "chz_batch_add": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]),
    "chz_batch_step": (ctypes.c_int, [ctypes.c_void_p]),
    "chz_batch_events": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Event), ctypes.c_int]),
    "chz_batch_live": (ctypes.c_int, [ctypes.c_void_p]),
    "chz_batch_nodes": (ctypes.c_longlong, [ctypes.c_void_p]),
    "chz_batch_set_budget": (None, [ctypes.c_void_p, ctypes.c_double]),
    "chz_batch_waits": (ctypes.c_longlong, [ctypes.c_void_p]),
    "chz_batch_frontier": (ctypes.c_longlong, [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_int)]),
}


# kernel families of chz_kernel_stats, in order (bit f of chz_time_kernels)
FAMILIES = ("expand", "compact", "sort", "dedupe", "gather", "level_graph")


def family_mask(families) -> int:
    """chz_time_kernels mask of family names; "all" for every family"""
    if families == "all":
        return (1 << len(FAMILIES)) - 1
    return sum(1 << FAMILIES.index(f) for f in families)


def lib():
    """the loaded HIP library; raises when it is missing (build it with __graft_entry__.build())"""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: the HIP extension is required (no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def level_line(lv: dict) -> str:
    """the reference's per-level stdout line (src/cheese.cu:1354-1356), with its "dir" prefix in the table and
    meet-in-the-middle modes, or the join line (:1714-1715)"""
    d = lv.get("dir", 0)
    if d == 4:
        return ('{"dir": "join", "forward": %d, "backward": %d, "hash_matches": %d, "seconds": %.2f}'
                % (lv["parents"], lv["prefixes"], lv["distinct"], lv["seconds"]))
    pre = '"dir": "%s", ' % DIRS[d] if d else ""
    return ('{%s"t": %d, "parents": %d, "prefixes": %d, "cut_w3": %d, "cut_lb": %d, "survivors": %d, '
            '"distinct": %d, "proved_unsat": %s, "seconds": %.2f}'
            % (pre, lv["t"], lv["parents"], lv["prefixes"], lv["cut_w3"], lv["cut_lb"], lv["survivors"],
               lv["distinct"], "true" if lv["proved_unsat"] else "false", lv["seconds"]))


class SearchError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"exit {rc}: {msg}")
        self.rc = rc


def search(spec_text: str, depth: int | None = None, dedupe: bool = True, sym: bool = True, chunk: int = 0,
           device: int = 0, stats: bool = False, table: int = 0, mitm: bool = False):
    """Runs the gate-count search on a .spec text (the reference's `cheese SPEC [flags]`; table=K is `--table K`,
    mitm `--mitm`).  Returns dict(rc, levels=[...], circuit=[names] or None, found, seconds, split, kernel_stats)."""
    L = lib()
    o = Opts(-1 if depth is None else depth, int(dedupe), int(sym), chunk, int(stats), device, table, int(mitm))
    ctx = ctypes.c_void_p()
    raw = spec_text.encode()
    rc = L.chz_open(raw, len(raw), ctypes.byref(o), ctypes.byref(ctx))
    if rc:
        return {"rc": rc, "levels": [], "circuit": None, "error": L.chz_last_error().decode()}
    levels = []

    def on_level(p, _user):
        v = p.contents
        levels.append({f: getattr(v, f) for f, _ in Level._fields_})

    cb = LEVEL_CB(on_level)
    res = Result()
    try:
        rc = L.chz_run(ctx, cb, None, ctypes.byref(res))
        out = {"rc": rc, "levels": levels, "circuit": None, "found": res.found, "seconds": res.seconds,
               "sym": L.chz_sym_size(ctx), "depth": L.chz_depth(ctx),
               "split": (("table", "mitm")[res.split_kind - 1], [res.split[0], res.split[1]]) if res.split_kind else None}
        if rc == 0:
            idx = [res.circuit[i] for i in range(res.len)]
            out["circuit_idx"] = idx
            out["circuit"] = [L.chz_gate_name(ctx, g).decode() for g in idx]
            arr = (ctypes.c_int * max(1, len(idx)))(*idx)
            out["verified"] = L.chz_verify(ctx, arr, len(idx)) == 1
        elif rc >= 2:
            out["error"] = L.chz_last_error().decode()
        if stats:
            ks = (KStat * 16)()
            nk = L.chz_kernel_stats(ctx, ks, 16)
            out["kernel_stats"] = {ks[i].name.decode(): {"launches": ks[i].launches, "ms": ks[i].ms, "work": ks[i].work}
                                   for i in range(nk)}
        return out
    finally:
        L.chz_close(ctx)


# expansion stage counts (chz_expand_stages)
STAGES = ("residuals", "structural", "cut_ranks", "t_bounds", "keys", "t_passes")


class Searcher:
    """one context reused for repeated searches of its spec (chz_open once, chz_run per search): device buffers,
    the stream and the tree pool persist between runs, as in a long-running caller"""

    def __init__(self, spec_text: str, depth: int | None = None, dedupe: bool = True, sym: bool = True,
                 chunk: int = 0, device: int = 0, table: int = 0, mitm: bool = False):
        L = lib()
        o = Opts(-1 if depth is None else depth, int(dedupe), int(sym), chunk, 0, device, table, int(mitm))
        self._ctx = ctypes.c_void_p()
        raw = spec_text.encode()
        rc = L.chz_open(raw, len(raw), ctypes.byref(o), ctypes.byref(self._ctx))
        if rc:
            raise SearchError(rc, L.chz_last_error().decode())

    def time_kernels(self, families="all") -> None:
        """kernel families whose launches record HIP events for run(stats=True)'s milliseconds (chz_time_kernels);
        none by default, since each timed launch costs host time"""
        lib().chz_time_kernels(self._ctx, family_mask(families))

    def tune(self, reps: int = 2) -> dict:
        """chz_tune: the fastest expansion schedule per level-size bucket, kept by this context"""
        buf = ctypes.create_string_buffer(1 << 14)
        rc = lib().chz_tune(self._ctx, reps, buf, len(buf))
        if rc:
            raise SearchError(rc, lib().chz_last_error().decode())
        return json.loads(buf.value.decode())

    def run(self, stats: bool = False):
        L = lib()
        levels = []

        def on_level(p, _user):
            v = p.contents
            levels.append({f: getattr(v, f) for f, _ in Level._fields_})

        cb = LEVEL_CB(on_level)
        res = Result()
        rc = L.chz_run(self._ctx, cb, None, ctypes.byref(res))
        out = {"rc": rc, "levels": levels, "found": res.found, "seconds": res.seconds,
               "circuit_idx": [res.circuit[i] for i in range(res.len)] if rc == 0 else None,
               "split": (("table", "mitm")[res.split_kind - 1], [res.split[0], res.split[1]]) if res.split_kind else None}
        if rc == 0:
            idx = out["circuit_idx"]
            out["circuit"] = [L.chz_gate_name(self._ctx, g).decode() for g in idx]
            arr = (ctypes.c_int * max(1, len(idx)))(*idx)
            out["verified"] = L.chz_verify(self._ctx, arr, len(idx)) == 1
        if stats:
            ks = (KStat * 16)()
            nk = L.chz_kernel_stats(self._ctx, ks, 16)
            out["kernel_stats"] = {ks[i].name.decode(): {"launches": ks[i].launches, "ms": ks[i].ms, "work": ks[i].work}
                                   for i in range(nk)}
            st = (ctypes.c_double * len(STAGES))()
            L.chz_expand_stages(self._ctx, st)
            out["expand_stages"] = dict(zip(STAGES, [int(v) for v in st]))
        return out

    def close(self):
        if getattr(self, "_ctx", None):
            lib().chz_close(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def result_line(out: dict) -> str:
    """the reference's result line (src/cheese.cu:1658-1662, :1800), with its "table" / "mitm" field, or the
    proof line of those modes (:1840, :1758)"""
    split = out.get("split")
    if out["rc"] == 1 and split:
        kind, (a, b) = split
        val = b if kind == "table" else "[%d, %d]" % (a, b)
        return '{"depth": %d, "proved_unsat": true, "%s": %s}' % (out["depth"], kind, val)
    if out["rc"] != 0:
        return ""
    if not out["circuit"]:
        return '{"found": 0, "circuit": []}'
    names = ", ".join(json.dumps(c) for c in out["circuit"])
    extra = ', "%s": [%d, %d]' % (split[0], split[1][0], split[1][1]) if split else ""
    return '{"t": %d, "found": %d, "circuit": [%s]%s, "seconds": %.2f}' % (
        len(out["circuit"]), out["found"] or 1, names, extra, out["seconds"])


def verify_spec(spec_text: str, circuit_idx) -> bool:
    """exact check of a candidate-index list against the spec's target, on the host (no device work)"""
    raw = spec_text.encode()
    arr = (ctypes.c_int * max(1, len(circuit_idx)))(*circuit_idx)
    return lib().chz_verify_spec(raw, len(raw), arr, len(circuit_idx)) == 1


class Batch:
    """searches that share a gate set advancing level by level in the same kernel launches (chz_batch_*).
    The context is opened on `base_spec` (any spec of that gate set); add() queues searches by tag."""

    def __init__(self, base_spec: str, device: int = 0, chunk: int = 0):
        L = lib()
        self._ctx = ctypes.c_void_p()
        raw = base_spec.encode()
        o = Opts(-1, 1, 1, chunk, 0, device, 0, 0)
        rc = L.chz_open(raw, len(raw), ctypes.byref(o), ctypes.byref(self._ctx))
        if rc:
            raise SearchError(rc, L.chz_last_error().decode())
        self.names = [L.chz_gate_name(self._ctx, g).decode() for g in range(L.chz_num_gates(self._ctx))]

    def add(self, spec_text: str, tag: int, depth: int | None = None) -> int:
        raw = spec_text.encode()
        return lib().chz_batch_add(self._ctx, raw, len(raw), -1 if depth is None else depth, tag)

    def step(self) -> int:
        return lib().chz_batch_step(self._ctx)

    def events(self):
        """queued events: (tag, "level", level dict) or (tag, "result", result dict)"""
        out = []
        buf = (Event * 256)()
        while True:
            n = lib().chz_batch_events(self._ctx, buf, 256)
            for i in range(n):
                e = buf[i]
                if e.kind == 0:
                    out.append((e.tag, "level", {f: getattr(e.level, f) for f, _ in Level._fields_}))
                else:
                    r = e.result
                    idx = [r.circuit[j] for j in range(r.len)] if r.rc == 0 else None
                    out.append((e.tag, "result", {"rc": r.rc, "circuit_idx": idx, "found": r.found,
                                                  "seconds": r.seconds,
                                                  "circuit": [self.names[g] for g in idx] if idx is not None else None}))
            if n < 256:
                return out

    @property
    def live(self) -> int:
        return lib().chz_batch_live(self._ctx)

    @property
    def nodes(self) -> int:
        return lib().chz_batch_nodes(self._ctx)

    def set_budget(self, nbytes: float):
        """per-level memory planning: searches that would not fit wait a step (chz_batch_set_budget)"""
        lib().chz_batch_set_budget(self._ctx, float(nbytes))

    def time_kernels(self, families="all") -> None:
        """kernel families whose launches record HIP events for kernel_stats' milliseconds (chz_time_kernels)"""
        lib().chz_time_kernels(self._ctx, family_mask(families))

    def kernel_stats(self) -> dict:
        """chz_kernel_stats of the batch's context: launches, milliseconds (of timed families) and work per kernel
        family"""
        ks = (KStat * 16)()
        nk = lib().chz_kernel_stats(self._ctx, ks, 16)
        return {ks[i].name.decode(): {"launches": ks[i].launches, "ms": ks[i].ms, "work": ks[i].work} for i in range(nk)}

    def frontier(self, max_nodes: int, stride: int = 1):
        """every stride-th node of the current frontier, at most max_nodes (chz_batch_frontier): a list of
        (tag, candidate indices from the root, bound word)"""
        paths = (ctypes.c_int * (max_nodes * DMAX))()
        lens = (ctypes.c_int * max_nodes)()
        bw = (ctypes.c_uint32 * max_nodes)()
        tags = (ctypes.c_int * max_nodes)()
        n = lib().chz_batch_frontier(self._ctx, max_nodes, stride, paths, lens, bw, tags)
        if n < 0:
            raise SearchError(5, lib().chz_last_error().decode())
        return [(tags[i], [paths[i * DMAX + t] for t in range(lens[i])], bw[i]) for i in range(n)]

    @property
    def waits(self) -> int:
        return lib().chz_batch_waits(self._ctx)

    def close(self):
        if getattr(self, "_ctx", None):
            lib().chz_close(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def search_many(specs, device: int = 0, budget: float = 0, stats: dict | None = None):
    """all searches in one batch (they must share a gate set); returns per spec the dict search() returns.  budget:
    per-level memory planning in bytes (0: none); stats (a dict) receives the number of search-levels that waited
    and the batch's kernel statistics"""
    b = Batch(specs[0], device)
    if stats is not None:
        b.time_kernels("all")
    if budget:
        b.set_budget(budget)
    outs = [{"rc": None, "levels": [], "circuit": None} for _ in specs]
    try:
        for i, sp in enumerate(specs):
            rc = b.add(sp, i)
            if rc:
                outs[i]["rc"] = rc
        while True:
            for tag, kind, d in b.events():
                o = outs[tag]
                if kind == "level":
                    o["levels"].append(d)
                else:
                    o.update(d)
                    if d["rc"] == 0:
                        o["verified"] = verify_spec(specs[tag], d["circuit_idx"])
            if not b.live:
                break
            rc = b.step()
            if rc:
                raise SearchError(rc, lib().chz_last_error().decode())
        if stats is not None:
            stats["kernel_stats"] = b.kernel_stats()
            stats["waits"] = b.waits
        return outs
    finally:
        b.close()
