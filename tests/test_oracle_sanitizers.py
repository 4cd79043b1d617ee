"""The threaded CPU oracle under ThreadSanitizer and AddressSanitizer/UBSan.

The oracle threads like the reference (static blocks of points or pixels,
a shared `running` flag that an out-of-bounds `stop` clears, per-thread
status aggregation, atomic out-of-bounds counters, per-thread adjoint
partial sums).  oracle/sanitize_check.cc drives every SIR and DAS entry
point with four threads on inputs that hit that shared state; this test
builds it with each sanitizer (oracle/sanitize.mk, g++, host code only) and
requires a clean run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def sanitized():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    subprocess.run(["make", "-s", "-C", ORACLE, "-f", "sanitize.mk"], check=True)
    return os.path.join(ORACLE, "_san")


@pytest.mark.parametrize("exe", ["tsan_check", "asan_check"])
def test_threaded_oracle_is_sanitizer_clean(sanitized, exe):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1")
    proc = subprocess.run([os.path.join(sanitized, exe)], capture_output=True, text=True,
                          timeout=300, env=env)
    assert proc.returncode == 0, proc.stderr[-4000:]
    assert "sanitize ok" in proc.stdout
    for tag in ("ThreadSanitizer", "AddressSanitizer", "runtime error", "LeakSanitizer"):
        assert tag not in proc.stderr, proc.stderr[-4000:]
