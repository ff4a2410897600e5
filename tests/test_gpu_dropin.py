"""The C++ drop-in API (include/kern_interp/*) through a C++ program that uses
it exactly as the reference's test suite does (tests/cpp/dropin_test.cpp),
one process per group of cases."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "kern-interp-old_amd", "_build", "dropin_test")


@pytest.mark.gpu
@pytest.mark.parametrize("group", ["api", "matrix", "matrix_big", "analytic", "fp32", "incremental"])
def test_cpp_dropin_api(group):
    print(f"[dropin] tests/cpp/dropin_test.cpp {group}", flush=True)
    assert os.path.exists(BIN), "build first (__graft_entry__.build())"
    r = subprocess.run([BIN, group], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout
