"""ctypes wrapper of the CPU ORACLE (oracle/_build/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
REF_DRIVER = os.path.join(_HERE, "_ref", "sdl_ref_driver")

GLIBC, CR = 0, 1
NUDGED = 2  # test only: CR made 0.4 % too large (a planted libm error)
STAGES = ("MD", "SG", "TP", "TL", "TCP", "TCL")
LAYER_CHAIN = 16  # stages value of run(): the layer-pair chain (its TCs are TCL objects)
WIDTH = {"MD": 2, "SG": 4, "TP": 8, "TL": 8, "TCP": 16, "TCL": 16}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_geom_load.restype = ctypes.c_void_p
        L.oracle_geom_load.argtypes = [ctypes.c_char_p] * 3
        L.oracle_geom_free.argtypes = [ctypes.c_void_p]
        L.oracle_run.restype = ctypes.c_void_p
        L.oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_int]
        L.oracle_run_sequence.restype = ctypes.c_void_p
        L.oracle_run_sequence.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        L.oracle_result_count.restype = ctypes.c_long
        L.oracle_result_count.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_result_copy.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.oracle_result_counters.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_result_layers.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_result_free.argtypes = [ctypes.c_void_p]
        L.oracle_var_count.restype = ctypes.c_int
        L.oracle_var_count.argtypes = [ctypes.c_int]
        L.oracle_result_vars.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.oracle_explain.restype = ctypes.c_int
        L.oracle_explain.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 5 + [ctypes.c_int] + [ctypes.c_void_p] * 2 + [ctypes.c_int, ctypes.c_void_p]
        L.oracle_module_decode.argtypes = [ctypes.c_uint, ctypes.c_void_p]
        L.oracle_kat_phi_mpi_pi.restype = ctypes.c_float
        L.oracle_kat_phi_mpi_pi.argtypes = [ctypes.c_float]
        L.oracle_kat_atan2.restype = ctypes.c_float
        L.oracle_kat_atan2.argtypes = [ctypes.c_float, ctypes.c_float]
        L.oracle_kat_center.argtypes = [ctypes.c_void_p] * 3
        L.oracle_kat_module_gap.restype = ctypes.c_float
        L.oracle_kat_module_gap.argtypes = [ctypes.c_uint, ctypes.c_void_p]
        L.oracle_kat_shift.argtypes = [ctypes.c_void_p, ctypes.c_uint] + [ctypes.c_void_p] * 3
        L.oracle_time.restype = ctypes.c_double
        L.oracle_time.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


_geoms = {}


def geometry(paths):
    key = tuple(paths)
    if key not in _geoms:
        h = lib().oracle_geom_load(*[p.encode() for p in paths])
        if not h:
            raise RuntimeError(f"oracle could not load geometry {paths}")
        _geoms[key] = h
    return _geoms[key]


def _ptrs(ev):
    return [ev.x.ctypes.data, ev.y.ctypes.data, ev.z.ctypes.data, ev.detId.ctypes.data, ev.idx.ctypes.data]


# Per-object variables (oracle_result_vars; the order of oracle/sdl_oracle.h)
# The keys of a stored tracklet's reco-variable map (Tracklet.cc:152-293 and the
# _v2 algorithm's own ten), in std::map order.
TL_KEYS = tuple("""
    algo alpha1GeV_OutLo betaAv betaAv_0th betaAv_1st betaAv_2nd betaAv_3rd betaAv_4th betaIn
    betaInRHmax betaInRHmin betaIn_0th betaIn_1st betaIn_1stCorr betaIn_2nd betaIn_2ndCorr
    betaIn_3rd betaIn_3rdCorr betaIn_4th betaIn_cut betaOut betaOutRHmax betaOutRHmin betaOut_0th
    betaOut_1st betaOut_1stCorr betaOut_2nd betaOut_2ndCorr betaOut_3rd betaOut_3rdCorr betaOut_4th
    betaPt_0th betaPt_1st betaPt_2nd betaPt_3rd betaPt_4th betacormode dBeta dBetaCut2 dBetaLum2
    dBetaMuls dBetaRIn2 dBetaROut2 dBetaRes dBeta_0th dBeta_1st dBeta_2nd dBeta_3rd dBeta_4th
    dBeta_midPoint dPhi deltaPhiPos deltaZLum dr drt_InSeg drt_tl_axis dzDrtScale hit1_x hit1_y
    hit2_x hit2_y hit3_x hit3_y hit4_x hit4_y innerSgInnerMdDetId innerSgOuterMdDetId k2Rinv1GeVf
    kRinv1GeVf kZ outerSgInnerMdDetId outerSgOuterMdDetId pixelPSZpitch ptCut pt_beta pt_betaIn
    pt_betaOut rawBetaIn rawBetaInCorrection rawBetaOut rawBetaOutCorrection rtHi_point rtIn
    rtLo_point rtOut rtOut_o_rtIn rtRatio_OutLoInLo rt_InLo rt_OutLo sdIn_alpha sdIn_d
    sdOut_alphaOut sdOut_d sdlCut sdlSlope sinAlphaMax strip2SZpitch zGeom zHi zHiPointed zIn zLo
    zLoPointed zOut z_InLo z_OutLo zpitch_InLo zpitch_OutLo
""".split())
VAR_NAMES = {
    "MD": ("dz", "shiftedDz", "deltaPhi", "deltaPhiChange", "deltaPhiNoShift", "deltaPhiChangeNoShift", "miniCut"),
    "SG": ("rtIn", "rtOut", "zIn", "zOut", "zLo", "zHi", "rtLo", "rtHi", "deltaPhi", "deltaPhiMin", "deltaPhiMax",
           "deltaPhiChange", "deltaPhiMinChange", "deltaPhiMaxChange", "sdCut", "sdSlope", "sdMuls", "sdPVoff",
           "reco_deltaPhi", "dAlpha_innerMD_segment", "dAlpha_outerMD_segment", "dAlpha_innerMD_outerMD"),
    "TL": ("betaIn", "betaInCut", "betaOut", "betaOutCut", "deltaBeta", "deltaBetaCut") + tuple(
        "reco_" + k for k in TL_KEYS),
    "TCL": ("dR", "innerR"),
}


def run(paths, ev, libm=GLIBC, stages=6, vars=False, layers=False):
    """Returns ({stage: int32 array [n, width]}, counters dict), and with
    vars=True also {"MD"|"SG"|"TL"|"TCL": float32 array [n, nvars]} (VAR_NAMES), the
    per-object variables in the same object order; with layers=True, last, the
    per-layer counters as an int64 array [10, 11] (counters in COUNTERS order x
    barrel layers 1-6, endcap layers 1-5).  stages = 1..6
    runs that many module-map stages; LAYER_CHAIN runs the layer-pair chain; a
    list runs those stages in that order (0-5 module-map, 6-8 createSegments,
    createTracklets, createTrackCandidates, 9 createTrackletsViaNavigation,
    10 createPseudoMiniDoubletsFromAnchorModule)."""
    L = lib()
    g = geometry(paths)
    if isinstance(stages, (list, tuple)):
        seq = np.asarray(stages, np.int32)
        r = L.oracle_run_sequence(g, ev.nhits, *_ptrs(ev), libm, seq.ctypes.data, len(seq))
    else:
        r = L.oracle_run(g, ev.nhits, *_ptrs(ev), libm, stages)
    try:
        out = {}
        for i, s in enumerate(STAGES):
            n = L.oracle_result_count(r, i)
            a = np.empty((n, WIDTH[s]), np.int32)
            if n:
                L.oracle_result_copy(r, i, a.ctypes.data)
            out[s] = a
        c = np.zeros(10, np.uint64)
        L.oracle_result_counters(r, c.ctypes.data)
        counters = dict(zip(("nMDCand", "nSGCand", "nTPCand", "nTLCand", "nTCCand", "nMD", "nSG", "nTP", "nTL", "nTC"),
                            [int(v) for v in c]))
        extra = ()
        if layers:
            lc = np.zeros(110, np.uint64)
            L.oracle_result_layers(r, lc.ctypes.data)
            extra = (lc.astype(np.int64).reshape(10, 11),)
        if not vars:
            return (out, counters) + extra
        v = {}
        for name, st in (("MD", 0), ("SG", 1), ("TL", 3), ("TCL", 5)):
            nv = L.oracle_var_count(st)
            a = np.zeros((len(out[name]), nv), np.float32)
            if len(a):
                L.oracle_result_vars(r, st, a.ctypes.data)
            v[name] = a
        return (out, counters, v) + extra
    finally:
        L.oracle_result_free(r)


def explain(paths, ev, kind, a, b=None, policy=None):
    """Re-evaluates one MD (kind 0, a = (lo, up)), SG (kind 1, a = 4 idx) or TL
    (kind 2, a, b = 8 idx) candidate under the GLIBC policy and `policy` (CR by
    default)."""
    L = lib()
    g = geometry(paths)
    aa = np.asarray(a, np.int32)
    bb = np.asarray(b if b is not None else [0] * 4, np.int32)
    out = np.zeros(9, np.float64)
    L.oracle_explain(g, ev.nhits, *_ptrs(ev), kind, aa.ctypes.data, bb.ctypes.data, CR if policy is None else policy,
                     out.ctypes.data)
    keys = ("cut_glibc", "cut_cr", "var", "lo", "hi", "var_cr", "lo_cr", "hi_cr", "ulps")
    return dict(zip(keys, out.tolist()))


def kat_center(xs, ys):
    """Centre of MathUtil::getCenterFromThreePoints through (xs[k], ys[k])."""
    x = np.asarray(xs, np.float32)
    y = np.asarray(ys, np.float32)
    out = np.zeros(2, np.float32)
    lib().oracle_kat_center(x.ctypes.data, y.ctypes.data, out.ctypes.data)
    return out


def kat_shift(paths, lower_det, lo, up):
    """MiniDoublet::shiftStripHits of hit lo on module lower_det and hit up on its partner."""
    a = np.asarray(lo, np.float32)
    b = np.asarray(up, np.float32)
    out = np.zeros(3, np.float32)
    lib().oracle_kat_shift(geometry(paths), lower_det, a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def time_events(paths, events, libm=GLIBC, threads=1):
    """Wall seconds to run the whole chain on `events` with `threads` threads."""
    L = lib()
    g = geometry(paths)
    off = np.zeros(len(events) + 1, np.int64)
    off[1:] = np.cumsum([e.nhits for e in events])
    cat = [np.ascontiguousarray(np.concatenate([getattr(e, f) for e in events])) for f in ("x", "y", "z", "detId", "idx")]
    return L.oracle_time(g, len(events), off.ctypes.data, *[c.ctypes.data for c in cat], libm, threads)


def parse_ref_dump(path):
    """Parses a reference-driver dump into {event: ({stage: array}, counters)}."""
    events = {}
    cur = None
    rows = None
    with open(path) as f:
        for line in f:
            parts = line.split()
            tag = parts[0]
            if tag == "E":
                cur = int(parts[1])
                rows = {s: [] for s in STAGES}
                events[cur] = [rows, {}]
            elif tag in WIDTH:
                rows[tag].append([int(v) for v in parts[1:]])
            elif tag == "N":
                events[cur][1].update(dict(zip(("nMDCand", "nSGCand", "nTPCand", "nTLCand", "nTCCand"), map(int, parts[1:]))))
            elif tag == "C":
                events[cur][1].update(dict(zip(("nMD", "nSG", "nTP", "nTL", "nTC"), map(int, parts[1:]))))
            elif tag == "L":  # per-layer counters: 10 kinds x (barrel 1-6, endcap 1-5)
                events[cur][1]["layers"] = np.asarray(list(map(int, parts[1:])), np.int64).reshape(10, 11)
    out = {}
    for e, (rows, cnt) in events.items():
        out[e] = ({s: np.asarray(rows[s], np.int32).reshape(-1, WIDTH[s]) for s in STAGES}, cnt)
    return out


def as_multiset(a):
    """Sorted rows: multiset comparison key for tuple arrays."""
    if len(a) == 0:
        return a
    return a[np.lexsort(a.T[::-1])]
