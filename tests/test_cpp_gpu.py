"""The reference's acceleration_error! and circular_orbit! tests, ported to the C++ interface
(tests/cpp/interaction_tests.cpp), run as a child process on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "interaction_tests")


def test_cpp_interface_reference_tests():
    assert os.path.exists(BIN), "build tests/cpp first (__graft_entry__.build)"
    r = subprocess.run([BIN, "100"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
