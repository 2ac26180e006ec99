"""Runs the C++ drop-in test programs (tests/cpp/*.cpp, built by
__graft_entry__.build()): the metric call operators and host point forms on the
CPU, the clustering cases on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def _run(name, *args):
    path = os.path.join(CPP, name)
    assert os.path.exists(path), f"tests/cpp/{name} is not built (run __graft_entry__.build())"
    r = subprocess.run([path, *args], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout


def test_cpp_metrics_and_host_points():
    _run("test_metrics")


@pytest.mark.gpu
def test_cpp_dropin_program():
    _run("test_clusterer", os.path.join(ROOT, "tests", "golden", "reference_data"))
