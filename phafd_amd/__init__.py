"""phafd_amd -- MI355X (gfx950) implementation of PHAFD's per-timestep hot path.

The compute lives in libphafd_hip.so (hand-written HIP kernels behind the C ABI
of include/phafd_hip.h); this package is its ctypes binding.
"""
from .api import *  # noqa: F401,F403
from .api import System, lib, LIB_PATH  # noqa: F401
