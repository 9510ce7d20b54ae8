// CPU ORACLE for the SDL segment-linking chain -- test infrastructure only.
//
// A from-scratch C++ restatement of the reference algorithm (bsathian/SDL),
// used by tests/ to check the HIP path, and by bench.py's cpu_baseline leg.
// Nothing in the product (sdl_amd/) may include, link or call it.
//
// Arithmetic follows the compiled reference exactly: float where the reference
// computes in float, double where its unqualified calls and double literals
// promote (sqrt, sin, tan, pow, atan2, 0.2*, 1./, M_PI, ...), and the
// integer-truncating abs(float) of libstdc++ where the reference calls the
// unqualified abs (MiniDoublet.cc:273,787, Segment.cc:743-744).
//
// Two libm policies:
//   LIBM_GLIBC : std::asin/atan/sin/cos/tan on floats call glibc's float
//                functions, as the compiled reference does;
//   LIBM_CR    : those float functions are evaluated in double and rounded
//                once to float (the correctly rounded value in practice),
//                which is what the HIP kernels compute.
// The GLIBC policy is pinned against the compiled reference; the CR policy is
// the bit-exact counterpart of the device path.  Differences between the two
// are libm rounding effects and are listed with their deciding cut margins.
#pragma once

#include <cstdint>

extern "C" {

enum { ORACLE_LIBM_GLIBC = 0, ORACLE_LIBM_CR = 1,
       // test only: the CR value made 0.4 % too large -- a planted libm
       // error that the tests of the difference listing must detect
       ORACLE_LIBM_NUDGED = 2 };

// Stage indices used by the result accessors.
enum { ORACLE_MD = 0, ORACLE_SG = 1, ORACLE_TP = 2, ORACLE_TL = 3, ORACLE_TCP = 4, ORACLE_TCL = 5 };

// Per-object variables (oracle_result_vars), in this order:
// MD (7): getDz, getShiftedDz, getDeltaPhi, getDeltaPhiChange, getDeltaPhiNoShift,
//         getDeltaPhiChangeNoShift, recovar miniCut (MiniDoublet.h);
// SG (22): getRtIn, getRtOut, getZIn, getZOut, getZLo, getZHi, getRtLo, getRtHi,
//         getDeltaPhi, getDeltaPhiMin, getDeltaPhiMax, getDeltaPhiChange,
//         getDeltaPhiMinChange, getDeltaPhiMaxChange, recovars sdCut, sdSlope,
//         sdMuls, sdPVoff, deltaPhi, dAlpha_innerMD_segment,
//         dAlpha_outerMD_segment, dAlpha_innerMD_outerMD (Segment.h);
// TL (114): getBetaIn, getBetaInCut, getBetaOut, getBetaOutCut, getDeltaBeta,
//         getDeltaBetaCut (Tracklet.h), then the 108 keys of the reco-variable
//         map (TrackletBase.cc:48-60) in std::map order, from "algo" to
//         "zpitch_OutLo" (oracle.py: VAR_NAMES); a key the TL's variant never
//         writes is NaN, one it resets and never sets is -999.
// TCL (2): recovars dR, innerR of a TC made from tracklets (TrackCandidate.cc:206-207).
enum { kMDVars = 7, kSGVars = 22, kTLTyped = 6, kTLKeys = 108, kTLVars = kTLTyped + kTLKeys, kTCLVars = 2 };

// The detId decoding of Module.cc:159-358: subdet, side, layer, rod, ring,
// module, isLower, isInverted, partnerDetId, moduleType (0 PS, 1 2S),
// moduleLayerType (0 Pixel, 1 Strip).
void oracle_module_decode(unsigned detId, unsigned out[11]);

// Single functions of the restatement, for known-answer checks:
// MathUtil::Phi_mpi_pi and ATan2 (MathUtil.cc:3-27); the centre (out2 = x, y)
// of MathUtil::getCenterFromThreePoints (MathUtil.cc:44-106) through
// (x[k], y[k]); MiniDoublet::moduleGapSize (returned) and
// isTighterTiltedModules (*tighter) of a module (MiniDoublet.cc:1242-1332);
// MiniDoublet::shiftStripHits (MiniDoublet.cc:989-1209) of a hit lo (x, y, z)
// on lower module lowerDetId and a hit up on its partner, with the geometry
// of oracle_geom_load and glibc's libm (out3 = shifted x, y, z).
float oracle_kat_phi_mpi_pi(float x);
float oracle_kat_atan2(float y, float x);
void oracle_kat_center(const float* x, const float* y, float* out2);
float oracle_kat_module_gap(unsigned detId, int* tighter);
void oracle_kat_shift(void* geom, unsigned lowerDetId, const float* lo, const float* up, float* out3);

// Geometry in the reference's three text formats (TiltedGeometry.cc:36,
// EndcapGeometry.cc:48, ModuleConnectionMap.cc:36-44).  Returns null on error.
void* oracle_geom_load(const char* tiltedPath, const char* endcapPath, const char* mapPath);
void oracle_geom_free(void* geom);

// stages value that runs the layer-pair chain instead: createMiniDoublets,
// createSegments, createTracklets, createTrackCandidates (Event.cc:245-255,
// 432-443, 599-609, 823-837); its TCs are reported as TCL objects.
enum { ORACLE_LAYER_CHAIN = 16 };

// Runs the first `stages` (1..6) module-map stages on one event (or, with
// ORACLE_LAYER_CHAIN, the layer-pair chain).
void* oracle_run(void* geom, int nhits, const float* x, const float* y, const float* z, const unsigned* detId,
                 const int* idx, int libm, int stages);
// Runs the stages seq[0..n-1] in that order (0-5 module-map stages, 6-8
// createSegments, createTracklets, createTrackCandidates, 9
// createTrackletsViaNavigation, 10 createPseudoMiniDoubletsFromAnchorModule);
// results as oracle_run.
void* oracle_run_sequence(void* geom, int nhits, const float* x, const float* y, const float* z,
                          const unsigned* detId, const int* idx, int libm, const int* seq, int n);
// Tuple width: MD 2, SG 4, TP 8, TL 8, TCP 16, TCL 16 hit indices.
long oracle_result_count(void* res, int stage);
void oracle_result_copy(void* res, int stage, int* out);
// Per-object variables of MDs, SGs and TLs (stage ORACLE_MD, ORACLE_SG,
// ORACLE_TL, ORACLE_TCL): oracle_var_count(stage) floats per object, in object order.
int oracle_var_count(int stage);
void oracle_result_vars(void* res, int stage, float* out);
// out[0..4] candidates (MD, SG, TP, TL, TC), out[5..9] objects (MD, SG, TP, TL, TC)
void oracle_result_counters(void* res, unsigned long long* out10);
// Per-layer counters: out110[11 * k + j], k = the ten counters above, j = 0-5
// barrel layers 1-6, 6-10 endcap layers 1-5 (Event.cc:1193-1344).
void oracle_result_layers(void* res, unsigned long long* out110);
void oracle_result_free(void* res);

// Re-evaluates one candidate under the GLIBC policy and under policy libmB and
// reports the cut where the decisions differ.  kind: 0 MD (a,b = lower/upper hit positions in the
// event arrays), 1 SG (a,b = MD tuples (4 ints) looked up in the GLIBC result),
// 2 TL (a,b = SG tuples (8 ints)).  out: [cutGlibc, cutB, var, lo, hi,
// var_b, lo_b, hi_b, ulpDistance].  Returns 0 on success.
int oracle_explain(void* geom, int nhits, const float* x, const float* y, const float* z, const unsigned* detId,
                   const int* idx, int kind, const int* a, const int* b, int libmB, double* out);

// Times `nevents` generated-elsewhere events run by `threads` threads (one event
// per thread at a time); events are passed as concatenated hit arrays with
// offsets[nevents+1].  Returns wall seconds.
double oracle_time(void* geom, int nevents, const long* offsets, const float* x, const float* y, const float* z,
                   const unsigned* detId, const int* idx, int libm, int threads);
}
