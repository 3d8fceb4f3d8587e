"""prism_amd: MI355X-native PRISM compress / progressive-retrieve path.

The compute path is libprism_amd.so (hand-written gfx950 HIP kernels behind
the C ABI in include/prism_amd.h).  This package loads that library, and
offers a thin Python orchestration of it (pipeline.py) used by the tests and
the benchmark.
"""
from ._lib import (PA_F32, PA_F64, PA_MODE_NB, PA_MODE_SM, Layout, PrismError, layout,  # noqa: F401
                   lib, LIB_PATH)

__all__ = ["PA_F32", "PA_F64", "PA_MODE_NB", "PA_MODE_SM", "Layout", "PrismError", "layout", "lib",
           "LIB_PATH"]
