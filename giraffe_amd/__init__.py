"""MI355X-native Newton-step hot path of GIRAFFE's static analysis.

The compute path is libgiraffe_amd.so (HIP kernels for gfx950 behind the C ABI
of include/giraffe_amd.h); `giraffe_amd.runtime` loads it and raises if it is
missing.  `giraffe_amd.model` builds model descriptions.
"""
__all__ = ["abi", "model", "runtime"]
