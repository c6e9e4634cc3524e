"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: tests/sanitize_driver.cpp runs every oracle
stage (PLOC with both key widths, conversion, traversal, pre-splits, reinsertion, refit, rebuilds, validation) on a
small seeded scene in a build of obvhs_oracle.cpp with -fsanitize=address,undefined and no recovery, so any
out-of-bounds access, use after free, leak or undefined operation fails the test."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(HERE, "..", "oracle", "obvhs_oracle.cpp")
DRIVER = os.path.join(HERE, "sanitize_driver.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_under_asan_and_ubsan():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "oracle_sanitized")
        # the oracle's own arithmetic flags (oracle/build_oracle.py) plus the sanitizers
        cmd = ["g++", "-O1", "-g", "-std=c++17", "-pthread", "-ffp-contract=off", "-fno-fast-math", "-msse2",
               "-mfpmath=sse", "-mno-fma", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
               "-fno-omit-frame-pointer", ORACLE, DRIVER, "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
        run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
        assert run.returncode == 0 and run.stdout.strip().endswith("ok"), (run.stdout[-2000:], run.stderr[-6000:])
