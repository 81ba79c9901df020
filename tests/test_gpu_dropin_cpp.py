"""Runs the C++ drop-in test program (tests/cpp/dropin_test.cpp): MTL4's own
test programs for this path, written against boost/numeric/{mtl,itl} headers
and executed on the GPU through libmtl4g.so."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "dropin_test")


def test_dropin_program_builds_against_headers():
    # CPU: the program is built (by __graft_entry__.build) and links the C ABI
    assert os.path.exists(EXE), "run __graft_entry__.build() first"


@pytest.mark.gpu
def test_dropin_cpp_program():
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(p.stdout[-4000:])
    passed = [l for l in p.stdout.splitlines() if l.startswith("PASS ")]
    failed = [l for l in p.stdout.splitlines() if l.startswith("FAIL ")]
    assert p.returncode == 0 and not failed, p.stdout[-4000:] + p.stderr[-2000:]
    assert len(passed) >= 16
