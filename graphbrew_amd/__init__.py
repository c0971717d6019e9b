"""MI355X-native GraphBrew hot path: GPU relabel + GAPBS kernels behind a C ABI.

graphbrew_amd.gbk     ctypes binding of libgbk.so (include/gbk.h)
graphbrew_amd.build   builds libgbk.so and the bench drivers with hipcc/g++
"""
__all__ = ["gbk", "build"]
