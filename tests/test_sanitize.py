"""Host code under the sanitizers, on the CPU: tests/sanitize/Makefile builds
tests/sanitize/host_checks.cpp with g++ twice, AddressSanitizer + UndefinedBehaviorSanitizer and
ThreadSanitizer, and runs both.  Covered: the exact threaded draw of the reference Gaussian starts
(MT19937 jump-ahead, gauss_mt.hpp) against the oracle's serial stream, the drivers' flag parsing
and n(mu) mixing (driver_host.hpp), and the CPU oracle."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("make") is None or shutil.which("g++") is None, reason="needs make and g++")
def test_host_code_clean_under_asan_ubsan_and_tsan():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "sanitize"), "run"], capture_output=True,
                       text=True, timeout=900)
    print(r.stdout[-4000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.count("host_checks ok") == 2
