"""libceos_gpu's host code under AddressSanitizer and UndefinedBehaviorSanitizer: the library built
with both sanitizers on its host pass (tests/sanitize/build_lib_asan.sh, about 5 minutes of
compilation, so not part of __graft_entry__.build(); the device code is the ordinary build) and
driven through its host paths
by tests/sanitize/lib_driver.cpp: CSR patterns, patch and gather maps, colours, launch schedules
and pipelines, both assembly modes, FP32 storage, the numerical tangent, Gauss-point state and
state files, BC tables, every preconditioner and Krylov method with the coarse-space queries, a
failing element, periodic contribution lists and an RVE batch (solve, subset solve, tile probe,
state file).  No invalid access or undefined behaviour may be reported.  (Leak checking is off:
the HIP runtime keeps allocations until the process ends.)"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "sanitize", "_build", "lib_driver")


def test_library_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("the sanitized driver is not built: run tests/sanitize/build_lib_asan.sh")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout[-3000:], flush=True)
    print(r.stderr[-6000:], flush=True)
    assert r.returncode == 0
    assert "lib driver: ok" in r.stdout
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
