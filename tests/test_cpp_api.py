"""Runs the C++ drop-in test program (tests/cpp/test_clusterer.cpp, built by
__graft_entry__.build()) on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_clusterer")


@pytest.mark.gpu
def test_cpp_dropin_program():
    assert os.path.exists(BIN), "tests/cpp/test_clusterer is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN, os.path.join(ROOT, "tests", "golden", "reference_data")],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout
