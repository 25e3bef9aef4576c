"""The library's host-side code and the CPU oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (no GPU): tests/sanitize/Makefile builds every
library source with the sanitizers on the host side only and the oracle
with the same LLVM clang; tests/sanitize/host_checks.cpp drives parameter
derivation, the tip fit, every file format, argument rejections and the
oracle's whole-run paths. Any finding aborts the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "_build", "sanitize", "host_checks")


@pytest.mark.skipif(not os.path.exists("/opt/rocm/llvm/bin/clang++"), reason="LLVM clang not found")
def test_host_code_and_oracle_are_clean_under_asan_and_ubsan(tmp_path):
    jobs = str(min(8, os.cpu_count() or 1))
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "sanitize"), "-j", jobs],
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert "host checks ok" in out
    shutil.rmtree(tmp_path, ignore_errors=True)
