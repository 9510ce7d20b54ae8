"""ctypes bindings of the synthetic detector and seeded event generator
(``sdl_amd/csrc/gen``).  Host-only C++; no GPU involved."""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libsdl_gen.so")
DATA_DIR = os.path.join(_HERE, "_data")

_lib = None
_lock = threading.Lock()
_geoms = {}


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(LIB_PATH)
        lib.sdlgen_geometry_new.restype = ctypes.c_void_p
        lib.sdlgen_geometry_new.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.sdlgen_geometry_write.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
        lib.sdlgen_geometry_nsensors.argtypes = [ctypes.c_void_p]
        lib.sdlgen_event_new.restype = ctypes.c_void_p
        lib.sdlgen_event_new.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_ulonglong]
        lib.sdlgen_event_nhits.argtypes = [ctypes.c_void_p]
        lib.sdlgen_event_ntracks.argtypes = [ctypes.c_void_p]
        lib.sdlgen_event_copy.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 6
        lib.sdlgen_event_free.argtypes = [ctypes.c_void_p]
        _lib = lib
    return _lib


def geom_kind(workload):
    """Workloads 1, 4 and 11 use the flat six-layer barrel; 2, 3, 5, 12 and 15
    the full tilted-barrel + endcap detector."""
    return 1 if workload in (1, 4, 11) else 2


MAP_TRACKS = {1: 50000, 2: 200000}


def geometry(kind):
    """Returns (handle, (tilted, endcap, map) file paths) for a detector kind,
    building it once per process and writing the three text files once per tree."""
    with _lock:
        if kind in _geoms:
            return _geoms[kind]
        lib = _load()
        h = lib.sdlgen_geometry_new(kind, MAP_TRACKS[kind])
        os.makedirs(DATA_DIR, exist_ok=True)
        paths = tuple(os.path.join(DATA_DIR, f"{n}_{kind}.txt") for n in ("tilted", "endcap", "map"))
        tmp = tuple(p + f".{os.getpid()}.tmp" for p in paths)
        lib.sdlgen_geometry_write(h, *[p.encode() for p in tmp])
        for t, p in zip(tmp, paths):
            os.replace(t, p)
        _geoms[kind] = (h, paths)
        return _geoms[kind]


class Event:
    """One generated event as flat host arrays (x, y, z float32; detId uint32;
    idx int32; simTrack int32, -1 for noise)."""

    def __init__(self, x, y, z, detId, idx, sim, ntracks):
        self.x, self.y, self.z, self.detId, self.idx, self.sim = x, y, z, detId, idx, sim
        self.ntracks = ntracks

    @property
    def nhits(self):
        return len(self.x)


def event(workload, ievent, seed=0):
    lib = _load()
    h, _ = geometry(geom_kind(workload))
    e = lib.sdlgen_event_new(h, workload, ievent, seed)
    try:
        n = lib.sdlgen_event_nhits(e)
        arrs = [np.empty(n, np.float32) for _ in range(3)] + [np.empty(n, np.uint32)] + [np.empty(n, np.int32) for _ in range(2)]
        lib.sdlgen_event_copy(e, *[a.ctypes.data for a in arrs])
        return Event(*arrs, lib.sdlgen_event_ntracks(e))
    finally:
        lib.sdlgen_event_free(e)
