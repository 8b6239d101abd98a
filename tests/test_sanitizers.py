"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the host code, run on the CPU: the
oracle (oracle/san_driver.c calls every entry point on small and edge-case inputs) and the C++
interface's host logic with the C ABI's argument checks (tests/cpp/host_tests.cpp, no GPU needed).
Both abort on the first sanitizer report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")

pytestmark = pytest.mark.skipif(shutil.which("make") is None, reason="needs make")


def _run(binary, env):
    r = subprocess.run([binary], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "runtime error" not in out and "AddressSanitizer" not in out, out
    return out


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    out = _run(os.path.join(ROOT, "oracle", "_san", "san_driver"), ENV)
    assert "ok (0 failures)" in out


def test_cpp_host_logic_under_sanitizers():
    lib = os.path.join(ROOT, "particular_amd", "libparticular_amd.so")
    if not os.path.exists(lib):
        pytest.skip("the HIP library is not built")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), "san"], check=True)
    # the HIP runtime's own allocations at exit are not this code's leaks
    env = dict(ENV, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    out = _run(os.path.join(ROOT, "tests", "cpp", "host_tests_san"), env)
    assert "ok (0 failures)" in out
