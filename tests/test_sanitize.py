"""The CPU oracle and the library's host-only sources (scene generator, PLY
and COLMAP readers) under AddressSanitizer and UndefinedBehaviorSanitizer:
tests/sanitize builds them with gcc -fsanitize=address,undefined
-fno-sanitize-recover=all and runs a driver over small scenes and valid,
truncated and oversized-count files (tests/sanitize/check.cpp). Any
sanitizer report fails the run. WGSL float -> int conversions saturate where
C++ leaves them undefined, so the oracle's conversions are the main target.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="gcc is missing")
def test_sanitizers_clean():
    b = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(HERE, "_build", "sanitize_check")], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "sanitize check ok" in r.stdout
