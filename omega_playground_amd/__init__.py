"""MI355X-native shallow-water hot path of omega_playground (gfx950 HIP, FP64).

The compute path is libomega_swe.so (HIP kernels behind the C ABI in
include/omega_swe.h); `swe` is its ctypes binding with the reference's class
names, `workloads` builds initial states.
"""
__all__ = ["swe", "workloads"]
