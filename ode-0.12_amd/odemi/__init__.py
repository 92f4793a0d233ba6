"""odemi -- Python harness over the MI355X-native ODE 0.12 step path.

The compute path is libodemi.so (HIP kernels for gfx950 behind a C ABI,
include/odemi.h); this package only loads it with ctypes, builds batch
descriptors from scene arrays and wraps the adjoint in torch.autograd.
Loading fails loudly when the library has not been built.
"""
from .capi import lib, load_library, LIB_PATH  # noqa: F401
from .batch import Batch, ORDER_REFERENCE, ORDER_COLORED  # noqa: F401
from . import scenes  # noqa: F401
