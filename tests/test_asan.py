"""Host sanitizer run: the CPU-side code (PGM I/O, synthetic generators,
schedule planner, oracle, CPU queue-phase model) built with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/asan/) and driven over ragged shapes,
malformed files and random cases.  CPU only."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "asan")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_host_code_under_sanitizers(tmp_path):
    b = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, timeout=600)
    assert b.returncode == 0, b.stderr.decode()[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(HERE, "asan_check"), str(tmp_path)], capture_output=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stderr.decode()[-4000:]
    assert b"all checks passed" in r.stdout
