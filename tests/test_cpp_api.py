"""Runs the C++ test runner of the lagrange::winding::FastWindingNumber drop-in class
(lagrange_amd/cpp/tests/test_fast_winding_number.cpp, built by __graft_entry__.build())."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "lagrange_amd", "cpp", "build", "test_fast_winding_number")


def test_cpp_runner_is_built():
    assert os.path.exists(BIN), "build with __graft_entry__.build()"


@pytest.mark.gpu
def test_cpp_drop_in_class():
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL PASSED" in r.stdout, r.stdout + r.stderr
