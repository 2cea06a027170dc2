"""The CPU oracle under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer (oracle/sanitize_main.cpp: detach, attach and target closeness with
collisions, and a falling cube, four simulations in four threads at once, as the tests
drive the oracle from a thread pool).  A leak, an out-of-bounds access, an undefined
operation or a data race fails the build target's run."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("make") is None or shutil.which("g++") is None, reason="needs make and g++")
def test_oracle_clean_under_asan_ubsan_and_tsan():
    r = subprocess.run(["make", "-C", ORACLE, "sanitize"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    for name in ("detach", "attach", "targets", "cube"):
        assert r.stdout.count(name + ":") == 2, r.stdout[-2000:]   # once per sanitizer build
    assert "ERROR: AddressSanitizer" not in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stdout
    assert "runtime error" not in r.stdout
