"""MI355X-native split-step Fourier GPE propagator and Sobolev minimiser.

The compute path is libatus_hip.so (HIP kernels for gfx950 behind the C-ABI
of include/atus_hip.h); `atus2_amd.hip` binds it for Python.
"""
__all__ = ["hip"]
