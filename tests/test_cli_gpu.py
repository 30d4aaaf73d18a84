"""The drop-in `cheese` binary (cheese_amd/bin/cheese, the reference binary's command line src/cheese.cu:1851-1895)
run as a process on golden specs: its stdout, minus the "seconds" fields, and its exit code equal the reference
CPU build's.  Needs an MI355X."""
import subprocess

import pytest

from conftest import GOLDEN, manifest, read_log, strip_seconds

import cheese_amd

pytestmark = pytest.mark.gpu

# proofs, found circuits (result line included), flag variants, the lite bound, the wide T transform, the table
# and join modes
CASES = ["cs-d4", "cs-d5", "qft_2-d6", "qft_2-d7", "toffoli-d9", "toffoli-d9-nosym", "toffoli-d8-nodedupe",
         "toffoli-d10-chunk1000", "ccz-d8", "relative-phase-toffoli-d9", "random3-s0", "random3-s5",
         "rus13-unitary-d12", "t-ladder-d42", "toffoli-d10-table5", "cs-d5-table3", "cs-d5-mitm", "toffoli-d9-mitm"]


@pytest.mark.parametrize("name", CASES)
def test_cli_stdout_and_exit_code(name):
    m = manifest()[name]
    print(f"case {name} flags {m['flags']}", flush=True)
    p = subprocess.run([cheese_amd.CLI_PATH, f"{GOLDEN}/specs/{name}.spec", *m["flags"]], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == m["rc"], p.stderr
    assert strip_seconds(p.stdout) == strip_seconds(read_log(name))
