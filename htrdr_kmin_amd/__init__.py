"""MI355X-native null-collision path-tracing core of htrdr-atmosphere.

The product is ``libhtgx.so`` (HIP kernels for gfx950 behind the C-ABI of
``include/htgx.h``); this package holds its sources (``csrc/``), the ctypes
declarations of that C-ABI (``_ffi``) and the host-side draw-map / image
writer mirror (``atmosphere``).
"""
from . import _ffi  # noqa: F401

__all__ = ["_ffi", "atmosphere"]
