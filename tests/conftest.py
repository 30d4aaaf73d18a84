import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP product path)")


def strip_seconds(text):
    """per-level / result lines without their wall-clock field"""
    out = []
    for line in text.strip().splitlines():
        d = json.loads(line)
        d.pop("seconds", None)
        out.append(d)
    return out


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def golden_cases(max_seconds=None):
    """names of fixtures that carry a reference log, optionally only those the reference ran quickly"""
    out = []
    for name, m in sorted(manifest().items()):
        if m.get("rc") is None:
            continue
        if max_seconds is not None and m.get("ref_seconds", 0) > max_seconds:
            continue
        out.append(name)
    return out


def read_spec(name):
    with open(os.path.join(GOLDEN, "specs", name + ".spec")) as f:
        return f.read()


def read_log(name):
    with open(os.path.join(GOLDEN, "logs", name + ".log")) as f:
        return f.read()


@pytest.fixture(scope="session")
def oracle_lib():
    """the oracle (test infrastructure) built from oracle/ if it is not there yet"""
    import ctypes
    path = os.path.join(ROOT, "oracle", "libcheese_oracle.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True, capture_output=True)
    L = ctypes.CDLL(path)
    L.orc_run.restype = ctypes.c_int
    L.orc_run.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
    L.orc_residual.restype = ctypes.c_int
    L.orc_residual.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_longlong)]
    L.orc_bound.restype = ctypes.c_int
    L.orc_bound.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int, ctypes.c_int]
    return L


def oracle_run(L, spec, depth=-1, dedupe=True, sym=True):
    import ctypes
    buf = ctypes.create_string_buffer(1 << 20)
    raw = spec.encode()
    rc = L.orc_run(raw, len(raw), depth, int(dedupe), int(sym), 0, buf, len(buf), None)
    return rc, buf.value.decode()


def flags_to_kwargs(flags):
    kw = {}
    i = 0
    while i < len(flags):
        f = flags[i]
        if f == "--no-sym":
            kw["sym"] = False
        elif f == "--no-dedupe":
            kw["dedupe"] = False
        elif f == "--chunk":
            kw["chunk"] = int(flags[i + 1])
            i += 1
        elif f == "--depth":
            kw["depth"] = int(flags[i + 1])
            i += 1
        elif f == "--table":
            kw["table"] = int(flags[i + 1])
            i += 1
        elif f == "--mitm":
            kw["mitm"] = True
        i += 1
    return kw


@pytest.fixture(scope="session")
def tc_oracle():
    """the T-count oracle (test infrastructure) built from oracle/ if it is not there yet"""
    import ctypes
    path = os.path.join(ROOT, "oracle", "libtcount_oracle.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True, capture_output=True)
    L = ctypes.CDLL(path)
    L.orc_tc_table.restype = ctypes.c_void_p
    L.orc_tc_table.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_long]
    L.orc_tc_table_size.restype = ctypes.c_long
    L.orc_tc_table_size.argtypes = [ctypes.c_void_p]
    L.orc_tc_search.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_long, ctypes.c_int, ctypes.c_char_p,
                                ctypes.c_long, ctypes.POINTER(ctypes.c_int)]
    L.orc_tc_free.argtypes = [ctypes.c_void_p]
    return L


def read_tspec(name):
    with open(os.path.join(GOLDEN, "tspec", name + ".tspec")) as f:
        return f.read()
