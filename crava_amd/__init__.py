"""MI355X-native CRAVA Fourier-domain Bayesian AVO inversion core.

The product is libcrava_hip.so (HIP kernels for gfx950 behind the C ABI in
include/crava_hip.h).  This package holds its build script, a ctypes binding,
thin torch-facing wrappers used by tests and benchmarks, and the seeded
synthetic-input generator.
"""
from ._lib import CravaHipError, lib  # noqa: F401

__all__ = ["CravaHipError", "lib"]
