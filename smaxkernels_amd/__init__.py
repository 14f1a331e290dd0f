"""smaxkernels_amd — MI355X (gfx950) HIP platform of the SmaxKernels SpMV path.

The product is libsmax_hip.so (C ABI in include/smax_hip.h) and the C++
SMAX::Interface mirror under smaxkernels_amd/src that drives it.  This Python
package holds the build recipe, a ctypes binding used by tests and the
benchmark, and the seeded workload generators.
"""
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
BUILD_DIR = PACKAGE_DIR / "_build"
HIP_LIB = BUILD_DIR / "libsmax_hip.so"

__all__ = ["PACKAGE_DIR", "BUILD_DIR", "HIP_LIB"]
