"""set_amd -- MI355X-native ideal-MHD finite-volume step (HIP, gfx950).

The compute path is set_amd/libsetmhd.so (HIP kernels behind a C ABI,
declared in include/setmhd.h).  set_amd.solver binds it with ctypes and
raises if the library is missing: there is no CPU fallback.
"""
__all__ = ["mesh", "solver"]
