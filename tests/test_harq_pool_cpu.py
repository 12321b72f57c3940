"""The serving engine's HARQ buffer pool (size classes over a coalescing
arena, processes by (cell, RNTI, process id), idle eviction), driven by a
C++ test program built here with AddressSanitizer and
UndefinedBehaviorSanitizer (host code only, no GPU)."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_harq_pool_under_sanitizers():
    src = os.path.join(ROOT, "tests", "native", "harq_pool_test.cpp")
    inc = os.path.join(ROOT, "aerial-cuda-accelerated-ran_amd", "csrc")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "harq_pool_test")
        cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
               "-fno-sanitize-recover=undefined", "-I" + inc, src, "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
                   UBSAN_OPTIONS="print_stacktrace=1")
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "all checks passed" in r.stdout
