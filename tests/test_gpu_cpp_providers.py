"""Runs the C++ provider test binary (tests/cpp/skbench_hip_tests.cpp), which
exercises the skbench-style C++20 classes over the C ABI on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "skbench_hip_tests")


def test_cpp_provider_suite(device):
    if not os.path.exists(BIN):
        pytest.fail("C++ test binary not built: run __graft_entry__.build()")
    print("[case] C++ provider suite", flush=True)
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    assert "all checks passed" in out.stdout
