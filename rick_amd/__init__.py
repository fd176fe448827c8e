"""MI355X-native (gfx950) implementation of RICK's imaging chain:
w-stacking gridding -> per-w-plane 2-D C2C FFT -> w-phase correction.

The compute lives in librick_hip.so (hand-written HIP kernels, C ABI in
include/rick.h); this package is its host-side interface.
"""
from .imaging import (FP32, FP64, GAUSS, GAUSS_HI_PRECISION, KAISERBESSEL, Plan, RickError,
                      fft_planes, phase_correction, plan_for, set_kernel, set_phase, wstack)

__all__ = ["Plan", "plan_for", "wstack", "fft_planes", "phase_correction", "set_kernel",
           "set_phase", "RickError", "FP64", "FP32", "GAUSS", "KAISERBESSEL",
           "GAUSS_HI_PRECISION"]
