"""The C ABI library loads on a machine without a GPU and exports every entry point that
include/ceos_gpu.h declares (no compute calls here)."""
import os
import re
import subprocess

from ceos_amd import gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(REPO, "include", "ceos_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ceos_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_expected_entry_points():
    names = declared_symbols()
    for required in ["ceos_mesh_create", "ceos_solve_constitutive", "ceos_internal_force",
                     "ceos_tangent", "ceos_csr_pattern", "ceos_update_mesh",
                     "ceos_switch_converged_state", "ceos_homogenize", "ceos_mesh_destroy"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    L = gpu.load_library()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(gpu.EXPORTS) == declared_symbols()


def test_exported_dynamic_symbols_match_header():
    out = subprocess.check_output(["nm", "-D", "--defined-only", gpu.LIB_PATH]).decode()
    exported = set(re.findall(r"\bT (ceos_[a-z0-9_]+)\b", out))
    assert set(declared_symbols()) <= exported
