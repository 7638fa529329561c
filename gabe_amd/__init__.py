"""MI355X-native lattice evolution for GABE (scalar fields + gravitational waves).

The compute path is libgabe_amd.so (hand-written gfx950 HIP kernels behind the
C ABI in include/gabe.h).  This package holds the library, its sources
(csrc/), the kept C++ host program (host/) and a ctypes binding (lattice.py).
"""
from .lattice import (GABE_FP32, GABE_FP64, GabeError, Lattice, load, numbins)  # noqa: F401
from .configs import (ModelConfig, lambda_phi4_config, reference_default_config)  # noqa: F401
