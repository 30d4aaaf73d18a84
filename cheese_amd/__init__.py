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
                ("distinct", ctypes.c_long), ("proved_unsat", ctypes.c_int), ("seconds", ctypes.c_double)]


class Opts(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_int), ("dedupe", ctypes.c_int), ("sym", ctypes.c_int), ("chunk", ctypes.c_long),
                ("profile", ctypes.c_int), ("device", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("rc", ctypes.c_int), ("len", ctypes.c_int), ("found", ctypes.c_long),
                ("circuit", ctypes.c_int * DMAX), ("seconds", ctypes.c_double)]


class KStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("launches", ctypes.c_long), ("ms", ctypes.c_double),
                ("work", ctypes.c_double)]


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
    "chz_verify": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
    "chz_last_error": (ctypes.c_char_p, []),
}


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
    """the reference's per-level stdout line (src/cheese.cu:1354-1356)"""
    return ('{"t": %d, "parents": %d, "prefixes": %d, "cut_w3": %d, "cut_lb": %d, "survivors": %d, '
            '"distinct": %d, "proved_unsat": %s, "seconds": %.2f}'
            % (lv["t"], lv["parents"], lv["prefixes"], lv["cut_w3"], lv["cut_lb"], lv["survivors"],
               lv["distinct"], "true" if lv["proved_unsat"] else "false", lv["seconds"]))


class SearchError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"exit {rc}: {msg}")
        self.rc = rc


def search(spec_text: str, depth: int | None = None, dedupe: bool = True, sym: bool = True, chunk: int = 0,
           device: int = 0, stats: bool = False):
    """Runs the gate-count search on a .spec text (the reference's `cheese SPEC [flags]`).
    Returns dict(rc, levels=[...], circuit=[names] or None, found, seconds, gates, kernel_stats)."""
    L = lib()
    o = Opts(-1 if depth is None else depth, int(dedupe), int(sym), chunk, 0, device)
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
               "sym": L.chz_sym_size(ctx)}
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


def result_line(out: dict) -> str:
    """the reference's result line (src/cheese.cu:1658-1662, :1800)"""
    if out["rc"] != 0:
        return ""
    if not out["circuit"]:
        return '{"found": 0, "circuit": []}'
    names = ", ".join(json.dumps(c) for c in out["circuit"])
    return '{"t": %d, "found": %d, "circuit": [%s], "seconds": %.2f}' % (
        len(out["circuit"]), out["found"] or 1, names, out["seconds"])
