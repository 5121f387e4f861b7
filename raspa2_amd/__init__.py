"""raspa2_amd -- MI355X-native engine for RASPA2's host-guest energy path.

The compute lives in librhg.so (HIP kernels for gfx950 behind the C ABI of
include/rhg.h); this package holds its ctypes binding (rhg), the synthetic
benchmark workloads (workloads) and the build script (build).
"""
from . import rhg  # noqa: F401  (fails loudly when librhg.so is missing)

__all__ = ["rhg"]
