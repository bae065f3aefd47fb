"""The frame-copy thread pool (qtrk_amd/csrc/copy_pool.hpp) that ScheduleFrame
uses to copy a frame's rows into pinned memory (the copy the reference makes
before ScheduleFrame returns, cputrack/QueuedTracker.cpp:91-111).

Host code only: the check program is built with g++ and ThreadSanitizer, so a
data race between the caller and the helper threads fails the test as well as
a wrong byte.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "qtrk_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_copy_pool_sizes_and_callers(tmp_path):
    exe = str(tmp_path / "copy_pool_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=thread", "-pthread", "-I", CSRC,
                    os.path.join(HERE, "copy_pool_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok" and "ThreadSanitizer" not in r.stderr
