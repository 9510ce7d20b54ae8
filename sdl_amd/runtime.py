"""ctypes binding of the C-ABI in include/sdl_amd.h (libsdl_amd.so).

Host plumbing only: every computation runs in the HIP kernels of the library.
Loading fails loudly when the library has not been built."""
import ctypes
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SDL_AMD_LIB selects another build of the same library (used to compare builds)
LIB_PATH = os.environ.get("SDL_AMD_LIB", os.path.join(_HERE, "_lib", "libsdl_amd.so"))

STAGE_MD, STAGE_SG, STAGE_TP, STAGE_TL, STAGE_TCP, STAGE_TCL = range(6)
STAGES = ("MD", "SG", "TP", "TL", "TCP", "TCL")
# layer-pair stages (createSegments, createTracklets, createTrackCandidates); their
# objects are exported as SG, TL and TCL
SG_LAYERS, TL_LAYERS, TC_LAYERS = 6, 7, 8
TL_NAV = 9  # createTrackletsViaNavigation; its objects are exported as TL
MD_PSEUDO = 10  # createPseudoMiniDoubletsFromAnchorModule; exported as MD
WIDTH = {"MD": 2, "SG": 4, "TP": 8, "TL": 8, "TCP": 16, "TCL": 16}
COUNTER_NAMES = ("nMDCand", "nSGCand", "nTPCand", "nTLCand", "nTCCand", "nMD", "nSG", "nTP", "nTL", "nTC")

# every entry point declared in include/sdl_amd.h
EXPORTS = (
    "sdl_last_error", "sdl_geom_load", "sdl_geom_free", "sdl_geom_num_modules", "sdl_event_create",
    "sdl_event_destroy", "sdl_event_reset", "sdl_event_add_hits", "sdl_event_next", "sdl_event_set_hits_device",
    "sdl_event_run", "sdl_event_run_chain", "sdl_event_sync", "sdl_event_counts", "sdl_event_layer_counts",
    "sdl_event_export", "sdl_event_export_noalloc", "sdl_object_var_count", "sdl_object_var_names", "sdl_event_export_vars",
    "sdl_event_num_events", "sdl_event_stream", "sdl_event_set_timing", "sdl_event_kernel_ms",
    "sdl_event_set_headroom", "sdl_event_set_graphs", "sdl_event_run_stats",
    "sdl_mlp_create", "sdl_mlp_free", "sdl_mlp_num_params", "sdl_mlp_set_params", "sdl_mlp_get_params",
    "sdl_mlp_gradient", "sdl_mlp_train", "sdl_mlp_forward", "sdl_event_set_proposal",
    "sdl_event_tl_pair_rows", "sdl_event_tl_work",
    "sdl_event_autotune", "sdl_event_get_tuning", "sdl_event_set_tuning",
    "sdl_geom_convert", "sdl_geom_save", "sdl_geom_load_binary", "sdl_geom_check_binary",
    "sdl_geom_load_tables", "sdl_module_decode", "sdl_map_build",
)
TUNING = ("md_block", "sg_block", "tl1_block", "tl2_block", "tl_expand_block", "scored_waves")
TIMERS = ("prep", "md", "sg", "tl")

_lib = None


class SDLError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SDLError(f"{LIB_PATH} is missing: the HIP library must be built (__graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
        L.sdl_last_error.restype = ctypes.c_char_p
        L.sdl_geom_load.argtypes = [ctypes.c_char_p] * 3 + [ctypes.POINTER(vp)]
        L.sdl_geom_free.argtypes = [vp]
        L.sdl_geom_num_modules.argtypes = [vp]
        L.sdl_event_create.argtypes = [vp, ctypes.POINTER(vp)]
        L.sdl_event_destroy.argtypes = [vp]
        L.sdl_event_reset.argtypes = [vp]
        L.sdl_event_add_hits.argtypes = [vp, i] + [vp] * 5
        L.sdl_event_next.argtypes = [vp]
        L.sdl_event_set_hits_device.argtypes = [vp, i, vp] + [vp] * 5
        L.sdl_event_run.argtypes = [vp, i, i]
        L.sdl_event_run_chain.argtypes = [vp]
        L.sdl_event_sync.argtypes = [vp]
        L.sdl_event_counts.argtypes = [vp, i, vp]
        L.sdl_event_layer_counts.argtypes = [vp, i, vp]
        L.sdl_event_export.argtypes = [vp, i, i, vp, ll, ctypes.POINTER(ll)]
        L.sdl_event_export_noalloc.argtypes = [vp, i, i, vp, ll, ctypes.POINTER(ll)]
        L.sdl_object_var_count.argtypes = [i]
        L.sdl_object_var_names.restype = ctypes.c_char_p
        L.sdl_object_var_names.argtypes = [i]
        L.sdl_event_export_vars.argtypes = [vp, i, i, vp, ll, ctypes.POINTER(ll)]
        L.sdl_event_num_events.argtypes = [vp]
        L.sdl_event_stream.argtypes = [vp]
        L.sdl_event_stream.restype = vp
        L.sdl_event_set_timing.argtypes = [vp, i]
        L.sdl_event_kernel_ms.argtypes = [vp, vp, i]
        L.sdl_event_set_headroom.argtypes = [vp, i, ctypes.c_double]
        L.sdl_event_set_graphs.argtypes = [vp, i]
        L.sdl_event_run_stats.argtypes = [vp, vp]
        f = ctypes.c_float
        L.sdl_mlp_create.argtypes = [ctypes.POINTER(vp)]
        L.sdl_mlp_free.argtypes = [vp]
        L.sdl_mlp_free.restype = None
        L.sdl_mlp_num_params.argtypes = []
        L.sdl_mlp_set_params.argtypes = [vp, vp, i]
        L.sdl_mlp_get_params.argtypes = [vp, vp, i]
        L.sdl_mlp_gradient.argtypes = [vp, vp, vp, ll, f, vp, vp]
        L.sdl_mlp_train.argtypes = [vp, vp, vp, vp, ll, i, i, f, f, vp]
        L.sdl_mlp_forward.argtypes = [vp, vp, ll, vp]
        L.sdl_event_set_proposal.argtypes = [vp, vp, f]
        L.sdl_event_tl_pair_rows.argtypes = [vp, vp, vp, ll, ctypes.POINTER(ll)]
        L.sdl_event_tl_work.argtypes = [vp, vp]
        L.sdl_event_autotune.argtypes = [vp, i, ctypes.c_char_p, i]
        L.sdl_geom_convert.argtypes = [ctypes.c_char_p] * 4
        L.sdl_geom_save.argtypes = [vp, ctypes.c_char_p]
        L.sdl_geom_load_binary.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.sdl_geom_check_binary.argtypes = [ctypes.c_char_p, ctypes.POINTER(i), ctypes.POINTER(ll)]
        L.sdl_geom_load_tables.argtypes = [i, vp, vp, vp, i, vp, vp, i, vp, vp, vp, ctypes.POINTER(vp)]
        L.sdl_module_decode.argtypes = [ctypes.c_uint, vp]
        L.sdl_map_build.argtypes = [ll, vp, vp, vp, ctypes.c_char_p, i]
        L.sdl_event_get_tuning.argtypes = [vp, vp, i]
        L.sdl_event_set_tuning.argtypes = [vp, vp, i]
        _lib = L
    return _lib


def var_names(stage):
    """Names of the per-object values of a stage (sdl_object_var_names)."""
    return tuple(lib().sdl_object_var_names(stage).decode().split(",")) if lib().sdl_object_var_count(stage) else ()


def _check(rc):
    if rc != 0:
        raise SDLError(f"sdl error {rc}: {lib().sdl_last_error().decode()}")


class Geometry:
    """The three geometry text files (reference formats), or with binary=PATH a
    binary geometry cache, loaded onto the current HIP device."""

    def __init__(self, tilted=None, endcap=None, cmap=None, binary=None):
        self.h = ctypes.c_void_p()
        if binary is not None:
            _check(lib().sdl_geom_load_binary(binary.encode(), ctypes.byref(self.h)))
        else:
            _check(lib().sdl_geom_load(tilted.encode(), endcap.encode(), cmap.encode(), ctypes.byref(self.h)))

    def save(self, path):
        """Writes the binary geometry cache of this geometry."""
        _check(lib().sdl_geom_save(self.h, path.encode()))

    @property
    def num_modules(self):
        return lib().sdl_geom_num_modules(self.h)

    def close(self):
        if self.h:
            lib().sdl_geom_free(self.h)
            self.h = ctypes.c_void_p()


def geometry_from_tables(tilted_rows, endcap_rows, conn):
    """A Geometry from tables in memory (sdl_geom_load_tables): tilted_rows
    [(detId, drdz, slope)], endcap_rows [(detId, 8 floats)], conn {detId: [detIds]}
    (in the order given)."""
    t = np.array([r[0] for r in tilted_rows], np.uint32), np.array([r[1] for r in tilted_rows], np.float32), \
        np.array([r[2] for r in tilted_rows], np.float32)
    ed = np.array([r[0] for r in endcap_rows], np.uint32)
    ec = np.array([list(r[1:]) for r in endcap_rows], np.float32).reshape(-1)
    md = np.array(list(conn.keys()), np.uint32)
    mc = np.array([len(v) for v in conn.values()], np.int32)
    ml = np.array([c for v in conn.values() for c in v], np.uint32)
    g = Geometry.__new__(Geometry)
    g.h = ctypes.c_void_p()
    _check(lib().sdl_geom_load_tables(len(t[0]), t[0].ctypes.data, t[1].ctypes.data, t[2].ctypes.data, len(ed),
                                      ed.ctypes.data, ec.ctypes.data, len(md), md.ctypes.data, mc.ctypes.data,
                                      ml.ctypes.data, ctypes.byref(g.h)))
    return g


MODULE_FIELDS = ("subdet", "side", "layer", "rod", "ring", "module", "isLower", "isInverted", "partnerDetId",
                 "moduleType", "moduleLayerType")


def decode_module(detId):
    """The reference's detId decoding (Module.cc:159-358) as a dict."""
    out = np.zeros(11, np.uint32)
    _check(lib().sdl_module_decode(int(detId), out.ctypes.data))
    return dict(zip(MODULE_FIELDS, (int(v) for v in out)))


def build_connection_map(tracks, path, merge=False):
    """Connection map from simulated tracks (sdl_map_build): tracks is a list of
    [(station, detId), ...] in station order; writes the reference's map format."""
    off = np.zeros(len(tracks) + 1, np.int64)
    off[1:] = np.cumsum([len(t) for t in tracks])
    st = np.array([s for t in tracks for s, _ in t], np.int32)
    det = np.array([d for t in tracks for _, d in t], np.uint32)
    _check(lib().sdl_map_build(len(tracks), off.ctypes.data, st.ctypes.data, det.ctypes.data, path.encode(),
                               1 if merge else 0))


def convert_geometry(tilted, endcap, cmap, binary):
    """Parses the three text files and writes the binary geometry cache (host only)."""
    _check(lib().sdl_geom_convert(tilted.encode(), endcap.encode(), cmap.encode(), binary.encode()))


def check_geometry_cache(binary):
    """Validates a binary geometry cache on the host: (modules, connections)."""
    nm, nc = ctypes.c_int(), ctypes.c_longlong()
    _check(lib().sdl_geom_check_binary(binary.encode(), ctypes.byref(nm), ctypes.byref(nc)))
    return nm.value, nc.value


class Event:
    """A batch of one or more events on one HIP stream."""

    def __init__(self, geom):
        self.geom = geom
        self.h = ctypes.c_void_p()
        _check(lib().sdl_event_create(geom.h, ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().sdl_event_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        _check(lib().sdl_event_reset(self.h))

    def add_hits(self, x, y, z, idx, detId):
        """Host arrays (float32 x3, int32, uint32), appended to the current event."""
        arrs = [np.ascontiguousarray(a, dt) for a, dt in
                ((x, np.float32), (y, np.float32), (z, np.float32), (idx, np.int32), (detId, np.uint32))]
        self._keep = arrs
        _check(lib().sdl_event_add_hits(self.h, len(arrs[0]), *[a.ctypes.data for a in arrs]))

    def next_event(self):
        _check(lib().sdl_event_next(self.h))

    def set_hits_device(self, offsets, x_ptr, y_ptr, z_ptr, idx_ptr, det_ptr):
        off = np.ascontiguousarray(offsets, np.int64)
        self._off = off
        _check(lib().sdl_event_set_hits_device(self.h, len(off) - 1, off.ctypes.data, x_ptr, y_ptr, z_ptr, idx_ptr, det_ptr))

    def run(self, stage, algo=0):
        _check(lib().sdl_event_run(self.h, stage, algo))

    def run_chain(self):
        _check(lib().sdl_event_run_chain(self.h))

    def sync(self):
        _check(lib().sdl_event_sync(self.h))

    @property
    def num_events(self):
        return lib().sdl_event_num_events(self.h)

    @property
    def stream(self):
        return lib().sdl_event_stream(self.h)

    def counts(self, e=-1):
        out = np.zeros(10, np.uint64)
        _check(lib().sdl_event_counts(self.h, e, out.ctypes.data))
        return dict(zip(COUNTER_NAMES, [int(v) for v in out]))

    def layer_counts(self, e=0):
        out = np.zeros(110, np.uint64)
        _check(lib().sdl_event_layer_counts(self.h, e, out.ctypes.data))
        return out.reshape(10, 11)

    def export(self, stage, e=0):
        name = STAGES[stage]
        n = ctypes.c_longlong(0)
        _check(lib().sdl_event_export(self.h, e, stage, None, 0, ctypes.byref(n)))
        out = np.empty((n.value, WIDTH[name]), np.int32)
        if n.value:
            _check(lib().sdl_event_export(self.h, e, stage, out.ctypes.data, out.size, ctypes.byref(n)))
        return out

    def export_vars(self, stage, e=0):
        """Per-object values of the MDs, SGs, TLs or TCs from tracklets of event e
        (float32 [n, nvars], object order of export; names: var_names(stage))."""
        L = lib()
        nv = L.sdl_object_var_count(stage)
        n = ctypes.c_longlong(0)
        _check(L.sdl_event_export_vars(self.h, e, stage, None, 0, ctypes.byref(n)))
        out = np.empty((n.value, nv), np.float32)
        if n.value:
            _check(L.sdl_event_export_vars(self.h, e, stage, out.ctypes.data, out.size, ctypes.byref(n)))
        return out

    def set_timing(self, on=True):
        _check(lib().sdl_event_set_timing(self.h, 1 if on else 0))

    def set_headroom(self, from_stage, factor):
        """Headroom factor of the predicted buffer sizes for stages >= from_stage."""
        _check(lib().sdl_event_set_headroom(self.h, from_stage, factor))

    def autotune(self, reps=3):
        """Times run_chain on the current batch for each launch-shape knob and
        keeps the fastest values; returns the report (timings and choice)."""
        buf = ctypes.create_string_buffer(4096)
        _check(lib().sdl_event_autotune(self.h, reps, buf, len(buf)))
        return json.loads(buf.value.decode())

    def tuning(self):
        v = np.zeros(len(TUNING), np.int32)
        rc = lib().sdl_event_get_tuning(self.h, v.ctypes.data, len(v))
        if rc != len(TUNING):
            _check(rc if rc < 0 else -1)
        return dict(zip(TUNING, v.tolist()))

    def set_tuning(self, t):
        v = np.asarray([t[k] for k in TUNING], np.int32)
        _check(lib().sdl_event_set_tuning(self.h, v.ctypes.data, len(v)))

    def tl_work(self):
        """(pairs, second-stage candidates, tracklets, pairs scored by the proposal)
        of the last tracklet stage."""
        out = np.zeros(4, np.int64)
        _check(lib().sdl_event_tl_work(self.h, out.ctypes.data))
        return dict(zip(("pairs", "stage2", "tracklets", "scored"), out.tolist()))

    def set_graphs(self, on=True):
        """Allow the speculative chain to be captured into a HIP graph and replayed."""
        _check(lib().sdl_event_set_graphs(self.h, 1 if on else 0))

    def run_stats(self):
        """Counts since the handle was created: stage runs, graph replays, graph
        captures and exact reruns (speculative sizes that fell short)."""
        out = np.zeros(4, np.int64)
        _check(lib().sdl_event_run_stats(self.h, out.ctypes.data))
        return dict(zip(("runs", "replays", "captures", "reruns"), (int(v) for v in out)))

    def kernel_ms(self):
        out = np.zeros(len(TIMERS), np.float64)
        _check(lib().sdl_event_kernel_ms(self.h, out.ctypes.data, len(TIMERS)))
        return dict(zip(TIMERS, out.tolist()))

    def export_all(self, e=0):
        return {s: self.export(i, e) for i, s in enumerate(STAGES)}
