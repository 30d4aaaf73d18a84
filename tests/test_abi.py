"""The C-ABI library loads and exports every entry point declared in include/cheese_hip.h; host-side
argument handling that needs no GPU.  CPU only."""
import os
import re
import subprocess

from conftest import ROOT

import cheese_amd


def _declared():
    with open(os.path.join(ROOT, "include", "cheese_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(chz_\w+)\s*\(", text)) - {"chz_level_cb"})


def test_library_exports_header_symbols():
    L = cheese_amd.lib()
    names = _declared()
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(L, n), n


def _cli(*args, stdin_spec=None, tmp_path=None):
    return subprocess.run([cheese_amd.CLI_PATH, *args], capture_output=True, text=True)


def test_cli_usage_exit_3():
    assert _cli().returncode == 3


def test_cli_bad_spec_exit_2(tmp_path):
    p = tmp_path / "bad.spec"
    p.write_text("cheese-spec 1\nn 2\nbogus 3\nend\n")
    assert _cli(str(p)).returncode == 2
    p.write_text("not a spec\n")
    assert _cli(str(p)).returncode == 2


def test_cli_unsupported_exit_3(tmp_path):
    spec = open(os.path.join(ROOT, "tests", "golden", "specs", "cs-d4.spec")).read()
    p = tmp_path / "s.spec"
    p.write_text(spec)
    assert _cli(str(p), "--depth", "49").returncode == 3          # depth > 48 (src/cheese.cu:1869)
    p.write_text(spec.replace("mode unitary", "mode states"))
    assert _cli(str(p)).returncode == 3                           # state targets are not built
    p.write_text(spec.replace("\nn 2\n", "\nn 4\n"))
    assert _cli(str(p)).returncode == 3                           # n outside 1..3


TCOUNT_CLI = os.path.join(ROOT, "cheese_amd", "bin", "tcount")


def test_tcount_cli_exit_codes(tmp_path):
    """the T-count command line's usage, bad-spec and unsupported exits (src/tcount.cu:436-457), before any
    device work"""
    assert subprocess.run([TCOUNT_CLI], capture_output=True).returncode == 3
    p = tmp_path / "bad.tspec"
    p.write_text("not a tspec\n")
    assert subprocess.run([TCOUNT_CLI, str(p)], capture_output=True).returncode == 2
    p.write_text("tcount-spec 1\nn 4\nend\n")
    assert subprocess.run([TCOUNT_CLI, str(p)], capture_output=True).returncode == 3
    assert subprocess.run([TCOUNT_CLI, str(tmp_path / "missing.tspec")], capture_output=True).returncode == 2
