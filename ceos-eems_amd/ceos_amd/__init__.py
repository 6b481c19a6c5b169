"""ceos_amd: MI355X-native element hot path of CEOS (hex8 constitutive update, F/K integration,
atomic-free CSR assembly, homogenization) behind the C ABI of include/ceos_gpu.h.

Import with the package folder on sys.path:  sys.path.insert(0, "<repo>/ceos-eems_amd").
"""
from . import meshgen  # noqa: F401

__all__ = ["meshgen", "gpu"]
