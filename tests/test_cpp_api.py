"""Runs the C++ test programs (tests/cpp/*.cpp, built by __graft_entry__.build()):
the metric call operators and host point forms, and the host-only runtime and the
oracle under the address and undefined-behaviour sanitizers, on the CPU; the
clustering cases of the drop-in on the GPU."""
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


def test_host_runtime_under_sanitizers():
    """Batch plans, workspace sizing, the staging ring and the server state machine
    (scripted device: normal runs, budget splits, failed launches, asynchronous
    device failures, back-pressure), built with ASan and UBSan."""
    _run("test_host_runtime")


def test_oracle_under_sanitizers():
    """Every oracle entry point over kernels, metrics, optional inputs, precisions and
    batched geometries, built with ASan and UBSan."""
    _run("test_oracle_sanitized")


def test_folded_gaussian_exp_equals_the_oracle():
    """exp((-d*d) * 2^-s) with the power of two folded into the exp's constants, as
    the float density stage evaluates the gaussian with 2*sigma*sigma = 2^s, gives
    the oracle's bits (every y of the core range for s = 1, a stride of all floats
    for |s| <= 8; `test_exp_scaled --exhaustive` widens both)."""
    _run("test_exp_scaled")


def test_sqrt_sign_selection_is_correctly_rounded():
    """The float square root of the neighbour stages settles the hardware estimate
    by the sign bits of two fma residuals added as integers: correctly rounded for
    every estimate within one ulp, on every float of [1, 4), at 0 and on a stride
    of all finite floats from 2^-96 up."""
    _run("test_sqrt_select")


@pytest.mark.gpu
def test_cpp_dropin_program():
    _run("test_clusterer", os.path.join(ROOT, "tests", "golden", "reference_data"))
