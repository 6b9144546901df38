"""gravidy_amd -- MI355X (gfx950) backend for GraviDy's 4th-order Hermite
block-timestep integrator.

The product is libh4hip.so (HIP kernels behind the C ABI in
include/h4hip.h) plus the C++ Hermite4HIP integrator and gravidy-hip
driver under host/.  This Python package is a thin ctypes view of the same
ABI, used by the tests and bench.py, and the initial-condition tools.
"""
from . import ic  # noqa: F401
from ._lib import (H4H_FP64, H4H_FP64_EXACT, H4H_MIXED, H4HError, LIB_PATH,  # noqa: F401
                   load)
from .ensemble import EnsembleHIP  # noqa: F401
from .hermite4 import (D_TIME_MAX, D_TIME_MIN, ETA_DEFAULT, Hermite4HIP,  # noqa: F401
                       softening_e2)
