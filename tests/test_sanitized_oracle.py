"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (gcc) and driven through
every entry point (tests/sanitize/oracle_driver.c: all material models, analytic and numerical
tangents, state commits, element-level calls, homogenization, a failing element): no invalid
access, leak or undefined behaviour is reported."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_oracle_is_clean_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "oracle_san"
    cmd = ["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fopenmp", "-ffp-contract=off", "-std=c99",
           os.path.join(HERE, "sanitize", "oracle_driver.c"), os.path.join(REPO, "oracle", "ceos_oracle.c"),
           "-lm", "-o", str(exe)]
    subprocess.check_call(cmd)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "oracle driver: ok" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
