"""velvet_amd -- MI355X-native XPBD cloth solver (HIP kernels for gfx950 behind a C ABI).

The compute path is velvet_amd/libvtx.so (include/vtx.h); this package holds its ctypes
binding (``solver``, with batching through segments), the seeded scene generators
(``scenes``), the serving driver (``server``) and the build script (``build``).
"""
from . import scenes  # noqa: F401

__all__ = ["scenes"]
