"""raspa2_amd -- MI355X-native engine for RASPA2's host-guest energy path.

The compute lives in librhg.so (HIP kernels for gfx950 behind the C ABI of
include/rhg.h).  Submodules: rhg (ctypes binding; importing it fails loudly
when librhg.so is missing), workloads (synthetic benchmark inputs), build
(hipcc build of librhg.so), serving (continuous-batching request loop).
"""
__all__ = ["rhg", "workloads", "build"]
