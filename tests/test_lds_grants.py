"""Dynamic-LDS limits requested from several host threads (CPU only).

The limit (hipFuncAttributeMaxDynamicSharedMemorySize) belongs to a kernel
for the whole process.  With a record of granted limits per thread, a thread
asking for less set its size over a larger one another thread had been
granted, and that thread's next launch asked for more dynamic LDS than the
limit in force.  The process-wide record
(kern-interp-old_amd/csrc/lds_grants.h) only ever raises a limit."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("lds") / "lds_grants_test")
    cmd = ["g++", "-std=c++17", "-O2", "-pthread", "-I",
           os.path.join(ROOT, "kern-interp-old_amd", "csrc"),
           os.path.join(ROOT, "tests", "cpp", "lds_grants_test.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def counts(exe, mode):
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    short, lowered = map(int, r.stdout.split())
    return short, lowered


def test_process_wide_record_never_lowers(exe):
    short, lowered = counts(exe, "process")
    assert lowered == 0 and short == 0


def test_per_thread_record_lowers_limits(exe):
    # the failure mode of a record per thread, shown on the CPU
    short, lowered = counts(exe, "thread")
    assert lowered > 0 and short > 0
