"""MI355X-native SDL segment-linking chain.

The product is a C-ABI shared library (``sdl_amd/_lib/libsdl_amd.so``, declared
in ``include/sdl_amd.h``) whose HIP kernels run the reference's module-map chain
(mini-doublets, segments, triplets, tracklets, track candidates) on one gfx950
GPU, and a C++ ``SDL::Event`` shim over it (``sdl_amd/include/SDL``).  This
Python package is host plumbing: ctypes bindings, the synthetic workload
generator and build helpers.
"""
from . import gen  # noqa: F401
