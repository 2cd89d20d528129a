"""AddressSanitizer + UndefinedBehaviorSanitizer on the CPU: the oracle
(oracle/nls_oracle.c) and libnls_amd's host-only code (operator builders,
BLAS-2 drop-ins, status reporting) are built instrumented by
tests/sanitize/Makefile and driven through every order, the smallest sizes
and the argument errors (tests/sanitize/*_check.*).  Any out-of-bounds
access, use after free, leak, overflow or other undefined behaviour aborts
the driver with a report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
OUT = os.path.join(ROOT, "_build", "sanitize")


@pytest.fixture(scope="module")
def built():
    if not shutil.which("gcc") or not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("needs gcc and hipcc")
    r = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    return OUT


@pytest.mark.parametrize("prog", ["oracle_check", "helpers_check"])
def test_clean_under_sanitizers(built, prog):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(built, prog)], capture_output=True, text=True, timeout=300, env=env)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "ERROR: AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert "%s: ok" % prog in report
