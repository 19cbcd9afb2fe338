"""MI355X (gfx950) fermion-inversion path of CL2QCD.

The kernels and the C ABI live in lib/libcl2qcd_hip.so (sources in csrc/,
declarations in include/cl2qcd.h); the C++ operator/solver drop-in lives in
host/.  `abi` binds the C ABI for Python callers.
"""
__all__ = ["abi"]
