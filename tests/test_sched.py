"""The render planner (crystal2_amd/csrc/c2_sched.h: samples -> chunks -> lanes) checked on
the CPU under AddressSanitizer and UndefinedBehaviorSanitizer for random single renders and
batches of uneven views, chunk sizes and 1..4 lanes (tests/sched_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_planner_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "sched_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "crystal2_amd", "csrc"),
                    os.path.join(HERE, "sched_check.cpp"), "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
