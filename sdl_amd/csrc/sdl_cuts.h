// Device-side cut algebra of the SDL chain for gfx950.
//
// Every function evaluates the reference's selection for ONE candidate in a
// single lane, in FP32 where the reference computes in float and in FP64 where
// its unqualified calls and double literals promote (A-16 of the reference's
// arithmetic: sqrt/sin/tan/pow/atan2 and 0.2*, 1./, M_PI terms).  The float
// transcendentals that the reference takes from libm in float (asinf, atanf,
// sinf, cosf, tanf) are evaluated in FP64 and rounded once, i.e. correctly
// rounded in practice; asin, atan, sin, cos, atan2 and a / sin(b) go through
// sdl_crmath.h, which returns the same float as that FP64 evaluation at a
// fraction of its cost.  The unqualified abs(float) of the reference binds to
// int abs(int) (MiniDoublet.cc:273,787, Segment.cc:743-744) and is emulated
// with the x86 truncation rule.  Compiled with -ffp-contract=off: no FMAs.
#pragma once

#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

#include "sdl_crmath.h"

namespace sdl {

constexpr double kPI = 3.14159265358979323846;
constexpr float kSlopeInfF = 123456789.0f;   // SDL_INF as compared in float (MiniDoublet.cc:3)
constexpr float kPtCut = 1.0f;               // PTCUT (Constants.h:4)
constexpr float kSinAlphaMax = 0.95f;
constexpr float kDeltaZLum = 15.f;
constexpr float kPixelPSZpitch = 0.15f;
constexpr float kStrip2SZpitch = 5.0f;

enum : uint8_t { SUB_ENDCAP = 4, SUB_BARREL = 5 };
enum : uint8_t { SIDE_NEGZ = 1, SIDE_POSZ = 2, SIDE_CENTER = 3 };
enum : uint8_t { TYPE_PS = 0, TYPE_2S = 1 };
enum : uint8_t { LT_PIXEL = 0, LT_STRIP = 1 };

// One sensor module of the geometry (32 bytes).  The strip-sensor lookups of
// the reference (drdz, slope) are resolved once at load time: for a lower
// module whose sensor is the pixel one they come from its partner
// (MiniDoublet.cc:779-781, 1032-1048; Segment.cc:926-929).
struct ModDev {
    uint32_t detId;
    int32_t partner;                 // module index of the partner sensor
    uint8_t subdet, side, layer, ring;
    uint8_t rod, type, layerType, isLower;
    float drdzStrip;                 // tiltedGeometry.getDrDz(strip sensor)
    float slopeStrip;                // barrel: tiltedGeometry.getSlope; endcap: endcapGeometry.getSlopeLower
    float centroidPhi;               // endcapGeometry.getCentroidPhi(own detId)
    float gap;                       // MiniDoublet::moduleGapSize
};

// k2Rinv1GeVf = float(float(2.99792458e-3 * 3.8) / 2.) (MiniDoublet.cc:746-747)
__device__ __forceinline__ float k2Rinv() {
    return float(double(float(2.99792458e-3 * 3.8)) / 2.);
}

__device__ __forceinline__ float asinF(float v) { return crm::asinCR(v); }    // float(asin(double(v)))
__device__ __forceinline__ float atanF(float v) { return crm::atanCR(v); }    // float(atan(double(v)))
__device__ __forceinline__ float sinF(float v) { return crm::sinCR(v); }      // float(sin(double(v)))
__device__ __forceinline__ float cosF(float v) { return crm::cosCR(v); }      // float(cos(double(v)))
__device__ __forceinline__ float tanF(float v) { return float(tan(double(v))); }
__device__ __forceinline__ float acosF(float v) { return float(acos(double(v))); }
__device__ __forceinline__ float sqrtF(float v) { return crm::sqrtCR(v); }      // float(sqrt(double(v)))
// std::abs(float) (the sign bit cleared, NaN included): a source modifier of
// the instruction that reads it
__device__ __forceinline__ float fabsF(float v) { return fabsf(v); }
// std::min / std::max argument-order semantics (NaN follows the first argument)
__device__ __forceinline__ float fmin2(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax2(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ double sq(double v) { return v * v; }

__device__ __forceinline__ int absTrunc(float v) {
    int i = (v > -2147483648.f && v < 2147483648.f) ? int(v) : INT_MIN;
    return i == INT_MIN ? INT_MIN : (i < 0 ? -i : i);
}

// MathUtil.cc:3-19.  For a float x, double(x) >= kPI iff x >= kPIUpF (the
// smallest float above pi) and double(x) < -kPI iff x <= -kPIUpF, so one float
// compare of |x| decides whether either loop runs (false for NaN as well).
constexpr float kPIUpF = 0x1.921fb6p+1f;
static_assert(double(kPIUpF) > kPI && double(0x1.921fb4p+1f) < kPI, "kPIUpF is the float above pi");
__device__ __forceinline__ float phiMpiPi(float x) {
    if (!(fabsf(x) >= kPIUpF) || isinf(x)) return x;
    while (double(x) >= kPI) x = float(double(x) - 2. * kPI);
    while (double(x) < -kPI) x = float(double(x) + 2. * kPI);
    return x;
}
// MathUtil.cc:21-27
__device__ __forceinline__ float atan2Sdl(float y, float x) {
    if (x != 0) return crm::atan2CR(y, x);  // float(atan2(double(y), double(x)))
    if (y == 0) return 0.f;
    if (y > 0) return float(kPI / 2);
    return float(-kPI / 2);
}
// derived quantities of a point (Hit.cc:68-83)
__device__ __forceinline__ float rtOf(float x, float y) { return sqrtF(x * x + y * y); }
__device__ __forceinline__ float r3Of(float x, float y, float z) { return sqrtF(x * x + y * y + z * z); }
__device__ __forceinline__ float phiOf(float x, float y) { return phiMpiPi(float(kPI + double(atan2Sdl(-y, -x)))); }
// Hit::deltaPhi (Hit.cc:149-152)
__device__ __forceinline__ float dPhi(float phiA, float phiB) { return phiMpiPi(phiB - phiA); }
// Hit::deltaPhiChange (Hit.cc:154-173): deltaPhi(a, b - a)
__device__ __forceinline__ float dPhiChange(float ax, float ay, float phiA, float bx, float by) {
    return phiMpiPi(phiOf(bx - ax, by - ay) - phiA);
}

struct HitView {
    float x, y, z, rt, r3, phi;
    float alpha;   // asin(min(rt k2 / ptCut, 0.95)), used as the MD threshold slope
};

// ------------------------------------------------------------------ MD
// MiniDoublet::moduleGapSize is precomputed per module (ModDev::gap).

// MiniDoublet::dPhiThreshold (MiniDoublet.cc:741-810)
// miniSlope = asin(min(rt k2 / ptCut, 0.95)) of the threshold hit is precomputed per hit.
__device__ __forceinline__ float mulsPtScaleBarrel(int iL) {
    return iL == 0 ? 0.0052f : iL == 1 ? 0.0038f : iL == 4 ? 0.0032f : 0.0034f;
}
__device__ __forceinline__ float rLayNominal(bool barrel, int iL) {
    if (barrel) return iL == 0 ? 21.8f : iL == 1 ? 34.6f : iL == 2 ? 49.6f : iL == 3 ? 67.4f : iL == 4 ? 87.6f : 106.8f;
    return iL == 0 ? 131.4f : iL == 1 ? 156.2f : iL == 2 ? 185.6f : iL == 3 ? 220.3f : 261.5f;
}
__device__ __forceinline__ float dPhiThreshold(float miniSlope, const ModDev& m, float dPhiArg, float dzArg) {
    const int iL = m.layer - 1;
    const bool barrel = m.subdet == SUB_BARREL;
    const float rLay = rLayNominal(barrel, iL);
    const float miniPVoff = float(0.1 / double(rLay));
    const float miniMuls = barrel ? mulsPtScaleBarrel(iL) * 3.f / 1.f : 0.006f * 3.f / 1.f;
    if (barrel && m.side == SIDE_CENTER) return float(double(miniSlope) + sqrt(sq(miniMuls) + sq(miniPVoff)));
    if (barrel) {
        const float d = m.drdzStrip;
        const float miniTilt = float(double(0.5f * kPixelPSZpitch * d) / sqrt(double(1.f + d * d)) / double(m.gap));
        return float(double(miniSlope) + sqrt(sq(miniMuls) + sq(miniPVoff) + sq(miniTilt * miniSlope)));
    }
    const float miniLum = float(absTrunc(dPhiArg * kDeltaZLum / dzArg));
    return float(double(miniSlope) + sqrt(sq(miniMuls) + sq(miniPVoff) + sq(miniLum)));
}

// MiniDoublet::shiftStripHits (MiniDoublet.cc:989-1209); pix / str are the
// pixel-side and strip-side hits of the pair.
__device__ __forceinline__ void shiftStripHits(const HitView& pix, const HitView& str, const ModDev& m, float& xn,
                                               float& yn, float& zn) {
    const bool isEndcap = m.subdet == SUB_ENDCAP;
    const float angleA = fabsF(atanF(pix.rt / pix.z));
    const float angleB = isEndcap ? float(kPI / 2.) : atanF(m.drdzStrip);
    float sep = m.gap;
    if (m.type == TYPE_PS && m.layerType != LT_PIXEL) sep *= -1;
    const float sAB = sinF(angleA + angleB);
    const float drprime = (sep / sAB) * sinF(angleA);
    const float slope = m.slopeStrip;
    const float absArctanSlope = slope != kSlopeInfF ? fabsF(atanF(slope)) : float(kPI / 2);
    const float xp = pix.x, yp = pix.y;
    float angleM;
    if (xp > 0 && yp > 0) angleM = absArctanSlope;
    else if (xp > 0 && yp < 0) angleM = float(kPI - double(absArctanSlope));
    else if (xp < 0 && yp < 0) angleM = float(kPI + double(absArctanSlope));
    else angleM = float(2 * kPI - double(absArctanSlope));
    const float xa = xp + drprime * sinF(angleM);
    const float ya = yp + drprime * cosF(angleM);
    const float xo = str.x, yo = str.y;
    if (slope == kSlopeInfF) { xn = xa; yn = yo; }
    else if (slope == 0) { xn = xo; yn = ya; }
    else {
        const float inv = 1.f / slope;
        xn = (slope * xa + inv * xo - ya + yo) / (slope + inv);
        yn = (xn - xa) * slope + ya;
    }
    const float absdzprime = fabsF(sep / sAB * cosF(angleA));
    const double az = double(fabsF(pix.z));
    const float abszn = m.layerType == LT_PIXEL ? float(az + double(absdzprime)) : float(az - double(absdzprime));
    zn = abszn * ((pix.z > 0) ? 1 : -1);
}

struct MDOut {
    float dz, shiftedDz, dphi, dphichange;
};

// runMiniDoubletDefaultAlgo{Barrel,Endcap} (MiniDoublet.cc:243-684).  rec (a
// passing MD's per-object values, kMDVars floats): getDz, getShiftedDz,
// getDeltaPhi, getDeltaPhiChange, the NoShift pair and the miniCut reco variable.
constexpr int kMDVars = 7, kSGVars = 22;
// TL: the six typed values (getBetaIn, getBetaInCut, getBetaOut, getBetaOutCut,
// getDeltaBeta, getDeltaBetaCut), then the 108 keys of the reco-variable map
// (TrackletBase.cc:48-60) in std::map order; a key the TL's variant never writes
// is NaN, one it resets (Tracklet.cc:152-248) and never sets is -999.
constexpr int kTLTyped = 6, kTLKeys = 108, kTLVars = kTLTyped + kTLKeys;
// TC from tracklets: recovars dR, innerR (TrackCandidate.cc:206-207).
constexpr int kTCLVars = 2;
enum TLKey {
    TK_algo, TK_alpha1GeV_OutLo, TK_betaAv, TK_betaAv_0th, TK_betaAv_1st, TK_betaAv_2nd, TK_betaAv_3rd,
    TK_betaAv_4th, TK_betaIn, TK_betaInRHmax, TK_betaInRHmin, TK_betaIn_0th, TK_betaIn_1st, TK_betaIn_1stCorr,
    TK_betaIn_2nd, TK_betaIn_2ndCorr, TK_betaIn_3rd, TK_betaIn_3rdCorr, TK_betaIn_4th, TK_betaIn_cut, TK_betaOut,
    TK_betaOutRHmax, TK_betaOutRHmin, TK_betaOut_0th, TK_betaOut_1st, TK_betaOut_1stCorr, TK_betaOut_2nd,
    TK_betaOut_2ndCorr, TK_betaOut_3rd, TK_betaOut_3rdCorr, TK_betaOut_4th, TK_betaPt_0th, TK_betaPt_1st,
    TK_betaPt_2nd, TK_betaPt_3rd, TK_betaPt_4th, TK_betacormode, TK_dBeta, TK_dBetaCut2, TK_dBetaLum2, TK_dBetaMuls,
    TK_dBetaRIn2, TK_dBetaROut2, TK_dBetaRes, TK_dBeta_0th, TK_dBeta_1st, TK_dBeta_2nd, TK_dBeta_3rd, TK_dBeta_4th,
    TK_dBeta_midPoint, TK_dPhi, TK_deltaPhiPos, TK_deltaZLum, TK_dr, TK_drt_InSeg, TK_drt_tl_axis, TK_dzDrtScale,
    TK_hit1_x, TK_hit1_y, TK_hit2_x, TK_hit2_y, TK_hit3_x, TK_hit3_y, TK_hit4_x, TK_hit4_y, TK_innerSgInnerMdDetId,
    TK_innerSgOuterMdDetId, TK_k2Rinv1GeVf, TK_kRinv1GeVf, TK_kZ, TK_outerSgInnerMdDetId, TK_outerSgOuterMdDetId,
    TK_pixelPSZpitch, TK_ptCut, TK_pt_beta, TK_pt_betaIn, TK_pt_betaOut, TK_rawBetaIn, TK_rawBetaInCorrection,
    TK_rawBetaOut, TK_rawBetaOutCorrection, TK_rtHi_point, TK_rtIn, TK_rtLo_point, TK_rtOut, TK_rtOut_o_rtIn,
    TK_rtRatio_OutLoInLo, TK_rt_InLo, TK_rt_OutLo, TK_sdIn_alpha, TK_sdIn_d, TK_sdOut_alphaOut, TK_sdOut_d,
    TK_sdlCut, TK_sdlSlope, TK_sinAlphaMax, TK_strip2SZpitch, TK_zGeom, TK_zHi, TK_zHiPointed, TK_zIn, TK_zLo,
    TK_zLoPointed, TK_zOut, TK_z_InLo, TK_z_OutLo, TK_zpitch_InLo, TK_zpitch_OutLo, kTLKeyCount
};
static_assert(int(kTLKeyCount) == kTLKeys, "tracklet key count");
__device__ __forceinline__ bool mdAlgo(const HitView& lo, const HitView& up, const ModDev& m, MDOut& o,
                                       float* rec = nullptr) {
    o.dz = lo.z - up.z;
    o.shiftedDz = 0.f;
    o.dphi = 0.f;
    o.dphichange = 0.f;
    const bool pixLower = m.layerType == LT_PIXEL;
    if (m.subdet == SUB_BARREL) {
        const float dz = o.dz;
        const float dzCut = m.type == TYPE_PS ? 2.f : 10.f;
        const float sign = float(((dz > 0) - (dz < 0)) * ((lo.z > 0) - (lo.z < 0)));
        const float invertedcrossercut = float(absTrunc(dz) > 2) * sign;
        if (!(fabsF(dz) < dzCut && invertedcrossercut <= 0)) return false;
        const float miniCut = pixLower ? dPhiThreshold(lo.alpha, m, 0.f, 1.f) : dPhiThreshold(up.alpha, m, 0.f, 1.f);
        if (m.side != SIDE_CENTER) {
            float xn, yn, zn;
            if (m.type == TYPE_PS) {
                if (pixLower) shiftStripHits(lo, up, m, xn, yn, zn);
                else shiftStripHits(up, lo, m, xn, yn, zn);
            } else {
                shiftStripHits(lo, up, m, xn, yn, zn);
            }
            const float sPhi = phiOf(xn, yn);
            const float sRt = rtOf(xn, yn);
            if (pixLower) o.dphi = dPhi(lo.phi, sPhi);
            else o.dphi = dPhi(sPhi, up.phi);
            if (!(fabsF(o.dphi) < miniCut)) return false;
            if (pixLower)
                o.dphichange = lo.rt < sRt ? dPhiChange(lo.x, lo.y, lo.phi, xn, yn) : dPhiChange(xn, yn, sPhi, lo.x, lo.y);
            else
                o.dphichange = sRt < up.rt ? dPhiChange(xn, yn, sPhi, up.x, up.y) : dPhiChange(up.x, up.y, up.phi, xn, yn);
        } else {
            o.dphi = dPhi(lo.phi, up.phi);
            if (!(fabsF(o.dphi) < miniCut)) return false;
            o.dphichange = dPhiChange(lo.x, lo.y, lo.phi, up.x, up.y);
        }
        if (rec) {  // MiniDoublet.cc:259, 327-343, 346, 391-410
            rec[0] = o.dz; rec[1] = o.shiftedDz; rec[2] = o.dphi; rec[3] = o.dphichange;
            rec[4] = dPhi(lo.phi, up.phi);
            rec[5] = (m.side == SIDE_CENTER || lo.rt < up.rt) ? dPhiChange(lo.x, lo.y, lo.phi, up.x, up.y)
                                                              : dPhiChange(up.x, up.y, up.phi, lo.x, lo.y);
            rec[6] = miniCut;
        }
        return fabsF(o.dphichange) < miniCut;
    }
    // endcap
    float dz = o.dz;
    if (!(fabsF(dz) < 10.f)) return false;  // side() is never Endcap (A-4)
    const float drtCut = m.type == TYPE_PS ? 2.f : 10.f;
    if (!(fabsF(lo.rt - up.rt) < drtCut)) return false;
    float xn, yn, zn;
    if (m.type == TYPE_PS && !pixLower) shiftStripHits(up, lo, m, xn, yn, zn);
    else shiftStripHits(lo, up, m, xn, yn, zn);
    const float sPhi = phiOf(xn, yn);
    if (m.type == TYPE_PS && !pixLower) o.dphi = dPhi(sPhi, up.phi);
    else o.dphi = dPhi(lo.phi, sPhi);
    if (m.type == TYPE_PS) {
        o.shiftedDz = pixLower ? lo.z - zn : up.z - zn;
        dz = o.shiftedDz;
    }
    const float miniCut = pixLower ? dPhiThreshold(lo.alpha, m, o.dphi, dz) : dPhiThreshold(up.alpha, m, o.dphi, dz);
    if (!(fabsF(o.dphi) < miniCut)) return false;
    const float dzFrac = float(double(fabsF(dz)) / double(fabsF(lo.z)));
    o.dphichange = o.dphi / dzFrac * (1.f + dzFrac);
    if (rec) {  // MiniDoublet.cc:467, 563-597, 608, 644-645
        rec[0] = o.dz; rec[1] = o.shiftedDz; rec[2] = o.dphi; rec[3] = o.dphichange;
        rec[4] = dPhi(lo.phi, up.phi);
        rec[5] = rec[4] / dzFrac * (1.f + dzFrac);
        rec[6] = miniCut;
    }
    return fabsF(o.dphichange) < miniCut;
}

// ------------------------------------------------------------------ SG
// What the segment algebra reads of a mini-doublet.  alpha1 = asin(min(rt k2, 0.95))
// and tanAlpha1 = tanf(alpha1) of the anchor are precomputed per MD.
struct MDView {
    HitView a;          // anchor hit
    float dz, shiftedDz, dphi, dphichange;
    float alpha1, tanAlpha1;
    int mod;            // lower module index
};

struct SGOut {
    float dphi, dphimin, dphimax, dphichange, dphichangemin, dphichangemax;
};

// Segment::dAlphaThreshold (Segment.cc:886-994) and the three alpha cuts
// (Segment.cc:538-592 / 789-853).
__device__ __forceinline__ bool sgAlphaCuts(const MDView& in, const MDView& out, const ModDev& mi, const ModDev& mo,
                                            float sgAlpha, float* rec = nullptr) {
    const bool inB = mi.subdet == SUB_BARREL, outB = mo.subdet == SUB_BARREL;
    const float sdMuls = inB ? mulsPtScaleBarrel(mi.layer - 1) * 3.f / kPtCut : 0.006f * 3.f / kPtCut;
    const float segmentDr = float(sqrt(sq(out.a.y - in.a.y) + sq(out.a.x - in.a.x)));
    const float dAlphaB = asinF(fmin2(segmentDr * k2Rinv() / kPtCut, kSinAlphaMax));
    float lumIn, lumOut;
    if (inB) {
        float t = 0.f;
        if (mi.side != SIDE_CENTER) {
            const float d = mi.drdzStrip;
            t = float(double(0.5f * kPixelPSZpitch * d) / sqrt(double(1.f + d * d)) / double(mi.gap));
        }
        lumIn = t * dAlphaB;
    } else {
        lumIn = float(double(in.dphi) * 15.0 / double(mi.type == TYPE_PS ? in.shiftedDz : in.dz));
    }
    if (outB) {
        float t = 0.f;
        if (mo.side != SIDE_CENTER) {
            const float d = mo.drdzStrip;
            t = float(double(0.5f * kPixelPSZpitch * d) / sqrt(double(1.f + d * d)) / double(mo.gap));
        }
        lumOut = t * dAlphaB;
    } else {
        lumOut = float(double(out.dphi) * 15.0 / double(mo.type == TYPE_PS ? out.shiftedDz : out.dz));
    }
    const float miniDelta = mi.gap;
    const float resIn = 0.02f / miniDelta * (inB ? 1.0f : fabsF(in.a.z / in.a.rt));
    const float resOut = 0.02f / miniDelta * (outB ? 1.0f : fabsF(out.a.z / out.a.rt));
    const float res = resIn + resOut;
    const double base = sq(res) + sq(sdMuls);
    const float thIO = float(double(dAlphaB) + sqrt(base));
    const float thIn = (inB && mi.side == SIDE_CENTER) ? thIO : float(double(dAlphaB) + sqrt(base + sq(lumIn)));
    const float thOut = (outB && mo.side == SIDE_CENTER) ? thIO : float(double(dAlphaB) + sqrt(base + sq(lumOut)));
    if (rec) {  // recovars dAlpha_* (Segment.cc:532-534, 801-803)
        rec[19] = thIn;
        rec[20] = thOut;
        rec[21] = thIO;
    }
    if (!(fabsF(in.dphichange - sgAlpha) < thIn)) return false;
    if (!(fabsF(out.dphichange - sgAlpha) < thOut)) return false;
    return fabsF(out.dphichange - in.dphichange) < thIO;
}

// runSegmentDefaultAlgo{Barrel,Endcap} (Segment.cc:337-859).  connFound: the
// outer module is in the connection list of the inner one (Segment.cc:627-634).
// rec (a passing SG's per-object values, kSGVars floats): rtIn, rtOut, zIn,
// zOut, zLo, zHi, rtLo, rtHi, the six deltaPhi values, then the reco variables
// sdCut, sdSlope, sdMuls, sdPVoff, deltaPhi and the three dAlpha thresholds;
// the typed fields a path never sets keep the constructor's 0, its unset reco
// variables -999 (Segment.cc:38-57, 382-390, 615-623).
__device__ __forceinline__ bool sgAlgo(const MDView& in, const MDView& out, const ModDev& mi, const ModDev& mo,
                                       bool connFound, const HitView& outHi, const HitView& outLo, SGOut& o,
                                       float* rec = nullptr) {
    o.dphi = o.dphimin = o.dphimax = o.dphichange = o.dphichangemin = o.dphichangemax = 0.f;
    const float rtIn = in.a.rt, rtOut = out.a.rt, zIn = in.a.z, zOut = out.a.z;
    const float sdSlope = out.alpha1;  // asin(min(rtOut k2 / ptCut, 0.95))
    if (mo.subdet == SUB_BARREL) {  // barrel-barrel, or endcap inner with barrel outer
        const float sdMuls = mulsPtScaleBarrel(mi.layer - 1) * 3.f / kPtCut;
        const float sdPVoff = 0.1f / rtOut;
        const float dzDrtScale = out.tanAlpha1 / sdSlope;
        const float zGeom = mi.layer <= 2 ? 2.f * kPixelPSZpitch : 2.f * kStrip2SZpitch;
        const float zLo = zIn + (zIn - kDeltaZLum) * (rtOut / rtIn - 1.f) * (zIn > 0.f ? 1.f : dzDrtScale) - zGeom;
        const float zHi = zIn + (zIn + kDeltaZLum) * (rtOut / rtIn - 1.f) * (zIn < 0.f ? 1.f : dzDrtScale) + zGeom;
        if (!(zOut >= zLo && zOut <= zHi)) return false;
        const float sdCut = float(double(sdSlope) + sqrt(double(sdMuls * sdMuls + sdPVoff * sdPVoff)));
        if (!(fabsF(dPhi(in.a.phi, out.a.phi)) <= sdCut)) return false;
        o.dphichange = dPhiChange(in.a.x, in.a.y, in.a.phi, out.a.x, out.a.y);
        if (rec) {  // Segment.cc:443-460, 480-484, 501
            const float v[19] = {rtIn, rtOut, zIn, zOut, zLo, zHi, 0.f, 0.f, 0.f, 0.f, 0.f, o.dphichange, 0.f, 0.f,
                                 sdCut, sdSlope, sdMuls, sdPVoff, dPhi(in.a.phi, out.a.phi)};
            for (int k = 0; k < 19; ++k) rec[k] = v[k];
        }
        if (!(fabsF(o.dphichange) <= sdCut)) return false;
        return sgAlphaCuts(in, out, mi, mo, o.dphichange, rec);
    }
    const bool l2r123 = (mi.ring == 1 || mi.ring == 2 || mi.ring == 3) && mi.layer == 2;
    if (!(connFound || l2r123)) return false;
    const float rtGeom = (rtIn < 60.f && rtOut < 60.f) ? 2.f * kPixelPSZpitch
                         : ((rtIn < 60.f || rtOut < 60.f) ? (kPixelPSZpitch + kStrip2SZpitch) : 2.f * kStrip2SZpitch);
    if (zIn * zOut < 0) return false;
    const float dz = zOut - zIn;
    const float dLum = copysignf(kDeltaZLum, zIn);
    const float drtDzScale = sdSlope / out.tanAlpha1;
    const float rtLo = fmax2(rtIn * (1.f + dz / (zIn + dLum) * drtDzScale) - rtGeom, rtIn - 0.5f * rtGeom);
    const float rtHi = rtIn * (zOut - dLum) / (zIn - dLum) + rtGeom;
    if (!(rtOut >= rtLo && rtOut <= rtHi)) return false;
    const float dPhiPos = dPhi(in.a.phi, out.a.phi);
    const float dHigh = dPhi(in.a.phi, outHi.phi);
    const float dLow = dPhi(in.a.phi, outLo.phi);
    const bool hiBigger = absTrunc(dHigh) > absTrunc(dLow);  // int abs (A-3)
    o.dphi = dPhiPos;
    o.dphimax = hiBigger ? dHigh : dLow;
    o.dphimin = hiBigger ? dLow : dHigh;
    if (!(fabsF(dPhiPos) <= sdSlope)) return false;
    const float dzFrac = dz / zIn;
    o.dphichange = dPhiPos / dzFrac * (1.f + dzFrac);
    o.dphichangemin = o.dphimin / dzFrac * (1.f + dzFrac);
    o.dphichangemax = o.dphimax / dzFrac * (1.f + dzFrac);
    if (rec) {  // Segment.cc:681-682, 721-722, 740-752, 769-771
        const float v[19] = {rtIn, rtOut, 0.f, 0.f, 0.f, 0.f, rtLo, rtHi, o.dphi, o.dphimin, o.dphimax, o.dphichange,
                             o.dphichangemin, o.dphichangemax, sdSlope, sdSlope, -999.f, -999.f, dPhiPos};
        for (int k = 0; k < 19; ++k) rec[k] = v[k];
    }
    if (!(fabsF(o.dphichange) <= sdSlope)) return false;
    return sgAlphaCuts(in, out, mi, mo, o.dphichange, rec);
}

// ------------------------------------------------------------------ TL
// Tracklet::runDeltaBetaIterations (Tracklet.cc:3181-3324): the values the cuts
// read and, with tk (a tracklet's map keys), the values it records.
__device__ __forceinline__ void tkIter(float* tk, int n, float bi, float bo, float ba, float bp) {
    // betaIn_<n>, betaOut_<n>, betaAv_<n>, betaPt_<n>, dBeta_<n> of iteration n
    const int in[5] = {TK_betaIn_0th, TK_betaIn_1st, TK_betaIn_2nd, TK_betaIn_3rd, TK_betaIn_4th};
    const int out[5] = {TK_betaOut_0th, TK_betaOut_1st, TK_betaOut_2nd, TK_betaOut_3rd, TK_betaOut_4th};
    const int av[5] = {TK_betaAv_0th, TK_betaAv_1st, TK_betaAv_2nd, TK_betaAv_3rd, TK_betaAv_4th};
    const int pt[5] = {TK_betaPt_0th, TK_betaPt_1st, TK_betaPt_2nd, TK_betaPt_3rd, TK_betaPt_4th};
    const int db[5] = {TK_dBeta_0th, TK_dBeta_1st, TK_dBeta_2nd, TK_dBeta_3rd, TK_dBeta_4th};
    tk[in[n]] = bi; tk[out[n]] = bo; tk[av[n]] = ba; tk[pt[n]] = bp; tk[db[n]] = bi - bo;
}

__device__ __forceinline__ void deltaBetaIter(float& betaIn, float& betaOut, float& betaAv, float& ptBeta, float sdInDr,
                                              float sdOutDr, float dr, int lIn, float* tk = nullptr) {
    const float ptBetaMax = 7.0f;
    const float k2 = k2Rinv();
    auto corr = [&](float len, float pt, float signOf) {
        return copysign(double(asinF(fmin2(len * k2 / fabsF(pt), kSinAlphaMax))), double(signOf));
    };
    auto ptOf = [&](float beta) { return crm::divSinCR(dr * k2, beta); };
    const bool mode1 = betaIn * betaOut > 0.f &&
                       (fabsF(ptBeta) < 4.f * ptBetaMax || (lIn >= 11 && fabsF(ptBeta) < 8.f * ptBetaMax));
    const bool mode2 = !mode1 && lIn < 11 && double(fabsF(betaOut)) < 0.2 * double(fabsF(betaIn)) &&
                       fabsF(ptBeta) < 12.f * ptBetaMax;
    if (!tk) {
        // The cuts' path: the two corrected modes below as one sequence (the same
        // operations on the same values), so that a wave with lanes in both
        // modes runs the corrections once.  Mode 2 starts from the pt of betaIn
        // and takes the sign of both corrections from betaIn (the updated one in
        // the second step); mode 1 from ptBeta, each correction with its own sign.
        if (!(mode1 || mode2)) return;
        float pt0 = ptBeta;
        if (mode2) pt0 = ptOf(betaIn);
        const float sign0 = mode1 ? betaOut : betaIn;
        const float inUpd = float(double(betaIn) + corr(sdInDr, pt0, betaIn));
        const float outUpd = float(double(betaOut) + corr(sdOutDr, pt0, sign0));
        betaAv = (mode1 || fabsF(betaOut) > 0.2f * fabsF(betaIn)) ? (0.5f * (inUpd + outUpd)) : inUpd;
        ptBeta = ptOf(betaAv);
        const float in2 = float(double(betaIn) + corr(sdInDr, ptBeta, betaIn));
        const float sign1 = mode1 ? betaOut : in2;
        betaOut = float(double(betaOut) + corr(sdOutDr, ptBeta, sign1));
        betaIn = in2;
        betaAv = 0.5f * (betaIn + betaOut);
        ptBeta = ptOf(betaAv);
        return;
    }
    if (mode1) {
        const float in0 = betaIn, out0 = betaOut;
        if (tk) {  // Tracklet.cc:3192-3200
            tk[TK_betacormode] = 1.f;
            tkIter(tk, 0, betaIn, betaOut, betaAv, ptBeta);
            tk[TK_betaIn_1stCorr] = float(corr(sdInDr, ptBeta, betaIn));
            tk[TK_betaOut_1stCorr] = float(corr(sdOutDr, ptBeta, betaOut));
        }
        const float inUpd = float(double(betaIn) + corr(sdInDr, ptBeta, betaIn));
        const float outUpd = float(double(betaOut) + corr(sdOutDr, ptBeta, betaOut));
        betaAv = 0.5f * (inUpd + outUpd);
        ptBeta = ptOf(betaAv);
        if (tk) {  // :3207-3213
            tkIter(tk, 1, inUpd, outUpd, betaAv, ptBeta);
            tk[TK_betaIn_2ndCorr] = float(corr(sdInDr, ptBeta, betaIn));
            tk[TK_betaOut_2ndCorr] = float(corr(sdOutDr, ptBeta, betaOut));
        }
        betaIn = float(double(betaIn) + corr(sdInDr, ptBeta, betaIn));
        betaOut = float(double(betaOut) + corr(sdOutDr, ptBeta, betaOut));
        betaAv = 0.5f * (betaIn + betaOut);
        ptBeta = ptOf(betaAv);
        if (tk) {  // :3221-3239: the 3rd restarts from the 0th values, the 4th from the 3rd
            tkIter(tk, 2, betaIn, betaOut, betaAv, ptBeta);
            tk[TK_betaIn_3rdCorr] = float(corr(sdInDr, ptBeta, betaIn));
            tk[TK_betaOut_3rdCorr] = float(corr(sdOutDr, ptBeta, betaOut));
            const float in3 = float(double(in0) + corr(sdInDr, ptBeta, betaIn));
            const float out3 = float(double(out0) + corr(sdOutDr, ptBeta, betaOut));
            const float av3 = 0.5f * (in3 + out3);
            const float pt3 = ptOf(av3);
            tkIter(tk, 3, in3, out3, av3, pt3);
            const float in4 = float(double(in0) + corr(sdInDr, pt3, in3));
            const float out4 = float(double(out0) + corr(sdOutDr, pt3, out3));
            const float av4 = 0.5f * (in4 + out4);
            tkIter(tk, 4, in4, out4, av4, ptOf(av4));
        }
    } else if (mode2) {
        if (tk) { tk[TK_betacormode] = 2.f; tkIter(tk, 0, betaIn, betaOut, betaAv, ptBeta); }  // :3245-3250
        const float ptBetaIn = ptOf(betaIn);
        const float inUpd = float(double(betaIn) + corr(sdInDr, ptBetaIn, betaIn));
        const float outUpd = float(double(betaOut) + corr(sdOutDr, ptBetaIn, betaIn));
        betaAv = (fabsF(betaOut) > 0.2f * fabsF(betaIn)) ? (0.5f * (inUpd + outUpd)) : inUpd;
        ptBeta = ptOf(betaAv);
        betaIn = float(double(betaIn) + corr(sdInDr, ptBeta, betaIn));
        betaOut = float(double(betaOut) + corr(sdOutDr, ptBeta, betaIn));
        betaAv = 0.5f * (betaIn + betaOut);
        ptBeta = ptOf(betaAv);
        if (tk)  // :3263-3285: the 1st ... 4th all hold the final values
            for (int n = 1; n < 5; ++n) tkIter(tk, n, betaIn, betaOut, betaAv, ptBeta);
    } else if (tk) {  // :3290-3320: nothing changes
        tk[TK_betacormode] = 3.f;
        for (int n = 0; n < 5; ++n) tkIter(tk, n, betaIn, betaOut, betaAv, ptBeta);
    }
}

// MathUtil::getCenterFromThreePoints (MathUtil.cc:44-106): the centre of the
// circle through A, B, C from the perpendicular bisectors of AB and BC, or the
// origin (a default Hit) when the two slopes are equal; returns the distance
// from A to it, as TrackCandidate.cc:196-197 computes it.
__device__ __forceinline__ float radiusFromA(float xA, float yA, float xB, float yB, float xC, float yC) {
    const float xMidAB = float(double(xA + xB) / 2.), yMidAB = float(double(yA + yB) / 2.);
    const bool infAB = (xB - xA) == 0, infBC = (xC - xB) == 0;
    const bool zeroAB = (yB - yA) == 0;
    const float sAB = infAB ? 0.f : (yB - yA) / (xB - xA);
    const float sBC = infBC ? 0.f : (yC - yB) / (xC - xB);
    const float perpAB = (infAB || zeroAB) ? 0.f : float(-1. / double(sAB));
    float cx = 0.f, cy = 0.f;
    if (!((sAB - sBC) == 0)) {
        const float num = sAB * sBC * (yA - yC) + sBC * (xA + xB) - sAB * (xB + xC);
        cx = float(double(num) / (2. * double(sBC - sAB)));
        cy = perpAB * (cx - xMidAB) + yMidAB;
    }
    const double dx = double(xA - cx), dy = double(yA - cy);
    return float(sqrt(dx * dx + dy * dy));
}

// The map keys the BBBB (_v2) and EEEE algorithms record alike at their end
// (Tracklet.cc:1399-1493, 2522-2617): the four anchors' x, y and module detIds
// (as float), the constants, the algorithm's own betacormode (0: :1245, :2351
// never change it) and the midpoint dBeta (:1409-1437, turning angles at the
// two segment midpoints).
__device__ __forceinline__ void tlCommonKeys(const HitView& h1, const HitView& h2, const HitView& h3, const HitView& h4,
                                             const uint32_t detId[4], float* tk) {
    tk[TK_hit1_x] = h1.x; tk[TK_hit1_y] = h1.y; tk[TK_hit2_x] = h2.x; tk[TK_hit2_y] = h2.y;
    tk[TK_hit3_x] = h3.x; tk[TK_hit3_y] = h3.y; tk[TK_hit4_x] = h4.x; tk[TK_hit4_y] = h4.y;
    tk[TK_innerSgInnerMdDetId] = float(detId[0]);
    tk[TK_innerSgOuterMdDetId] = float(detId[1]);
    tk[TK_outerSgInnerMdDetId] = float(detId[2]);
    tk[TK_outerSgOuterMdDetId] = float(detId[3]);
    tk[TK_sinAlphaMax] = kSinAlphaMax;
    tk[TK_deltaZLum] = kDeltaZLum;
    tk[TK_k2Rinv1GeVf] = k2Rinv();
    tk[TK_kRinv1GeVf] = float(2.99792458e-3 * 3.8);
    tk[TK_pixelPSZpitch] = kPixelPSZpitch;
    tk[TK_strip2SZpitch] = kStrip2SZpitch;
    tk[TK_ptCut] = kPtCut;
    tk[TK_betacormode] = 0.f;
    const float inMidX = float(double(h1.x + h2.x) / 2.), inMidY = float(double(h1.y + h2.y) / 2.);
    const float outMidX = float(double(h3.x + h4.x) / 2.), outMidY = float(double(h3.y + h4.y) / 2.);
    const float ax = h2.x - inMidX, ay = h2.y - inMidY;
    const float bx = outMidX - inMidX, by = outMidY - inMidY;
    const float cx = h4.x - outMidX, cy = h4.y - outMidY;
    const float aMag = sqrtF(ax * ax + ay * ay), bMag = sqrtF(bx * bx + by * by), cMag = sqrtF(cx * cx + cy * cy);
    const float ab = ax * bx + ay * by, bc = bx * cx + by * cy;
    tk[TK_dBeta_midPoint] = acosF(ab / aMag / bMag) - acosF(bc / bMag / cMag);
}

__device__ __forceinline__ void rescaleRH(float beta, float& mn, float& mx) {
    const float s = mn + mx;
    const float sf = (fabsF(s) > 0) ? (2.f * beta / fabsF(s)) : 0.f;
    mn *= sf;
    mx *= sf;
}

// Hit() += a, += b, /= 2 (Tracklet.cc:1131-1134) -> phi of the midpoint
__device__ __forceinline__ float midPhi(const HitView& a, const HitView& b) {
    const float mx = ((0.f + a.x) + b.x) / 2.0f, my = ((0.f + a.y) + b.y) / 2.0f;
    return phiOf(mx, my);
}

// Edge hit of GeometryUtil::stripEdgeHit (GeometryUtil.cc:3-23)
__device__ __forceinline__ HitView edgeOf(const HitView& h, float centroidPhi, bool up) {
    const double s = up ? 1.0 : -1.0;
    HitView e;
    e.alpha = 0.f;
    e.x = float(s * 2.5 * cos(double(centroidPhi))) + h.x;
    e.y = float(s * 2.5 * sin(double(centroidPhi))) + h.y;
    e.z = 0.f + h.z;
    e.rt = rtOf(e.x, e.y);
    e.r3 = r3Of(e.x, e.y, e.z);
    e.phi = phiOf(e.x, e.y);
    return e;
}

// --- Tracklets.  The kernel evaluates a candidate in two stages: the cuts that
// need only the inner segment (wave-uniform) and the outer segment's inner MD
// run first; the outer segment's outer MD and the per-segment quantities are
// loaded only for candidates that survive them.  All cuts must pass, so their
// order does not change the decision.  Quantities that depend on a single MD
// or segment are precomputed once per object with the same expressions:
//   MD:  alpha1 = asin(min(rt k2, 0.95)), dzScale = tan(alpha1)/alpha1 (double),
//        alphaReg3 = asin(min(rt k2 / 3, 0.95))               (anchor rt)
//   SG:  dr = rt(out - in), d = out.rt - in.rt, aOU = deltaPhi(out, out - in),
//        dPhiPos = deltaPhi(in, out), and aOU with the strip edges of out.

// wave-uniform view of the inner segment
struct TLInner {
    HitView lo, up;
    float dphichange, dphichangemin, dphichangemax;
    float dr, d, alphaReg3Lo;
    double sqrtR3Rt;  // sqrt(double(sqrtf(z^2 + rt^2) / rt)) of the inner anchor (the muls factor)
    bool bLo, bUp, psLo, psUp;
};
// outer segment, first stage: its inner MD
struct TLOutLo {
    HitView a;
    float alpha1, dzScale, alphaReg3;
    bool barrel, ps;
};
// outer segment, second stage
struct TLOutUp {
    HitView a;
    float dphichange, dphichangemin, dphichangemax, dphimin, dphimax;
    float dr, d, aOU, dPhiPos;
    bool ec2s;
    float hiX, hiY, hiRt, hiPhi, loX, loY, loRt, loPhi, aOUhi, aOUlo;
};

// Values of the first stage that the second stage reads.
struct TLCarry {
    float sdlCut, thetaMuls, sinDPhiMid;  // sin of the midpoint deltaPhi (BBBB, BBEE)
};

// Where a pair that passed the first stage lies inside its windows (features of
// the learned proposal): w1 = position in the z (BBBB) or rt (BBEE, EEEE)
// window, w2 = position in the pointed window (0.5 where the variant has none),
// both in [0, 1]; phi = midpoint deltaPhi / sdlCut, in [-1, 1].
struct TLMargins {
    float w1, w2, phi;
};

enum { TL_BBBB = 0, TL_BBEE = 1, TL_EEEE = 2, TL_NONE = 3 };

// Variant of runTrackletDefaultAlgo (Tracklet.cc:144-296) from the first three anchor
// subdets; the fourth (outer segment's outer MD) is checked in the second stage:
// BBBB and BBBE run _v2, BBEE and BEEE the barrel-endcap algorithm, EEEE the endcap
// one, every other combination fails.
__device__ __forceinline__ int tlVariant(bool b0, bool b1, bool b2) {
    if (b0 && b1 && b2) return TL_BBBB;
    if (b0 && !b2) return TL_BBEE;
    if (!b0 && !b1 && !b2) return TL_EEEE;
    return TL_NONE;
}

// sqrt(double(r / rt)) with r = sqrtf(z^2 + rt^2) of an anchor: the factor of
// the multiple-scattering term that depends on the inner anchor only
// (Tracklet.cc:1050, 1716)
__device__ __forceinline__ double mulsFactor(float z, float rt) {
    const float r = sqrtF(z * z + rt * rt);
    return sqrt(double(r / rt));
}

// ---- first stage: cuts on the inner segment and the outer segment's inner MD
// _v2: Tracklet.cc:902-1153 (z window, pointed z window, deltaPhi of the midpoint)
// tk (a stored TL's map keys, diagnostics only): the window values it records.
__device__ __forceinline__ bool tlStage1BBBB(const TLInner& I, const TLOutLo& L, TLCarry& c, TLMargins* m,
                                             float* tk = nullptr) {
    const HitView& hInLo = I.lo;
    const HitView& hInUp = I.up;
    const HitView& hOutLo = L.a;
    const float rt_InLo = hInLo.rt, rt_OutLo = hOutLo.rt, z_InLo = hInLo.z, z_OutLo = hOutLo.z;
    const float rtRatio = rt_OutLo / rt_InLo;
    const float dzDrtScale = L.dzScale;
    const float zpIn = I.psLo ? kPixelPSZpitch : kStrip2SZpitch;
    const float zpOut = L.ps ? kPixelPSZpitch : kStrip2SZpitch;
    const float zHi = z_InLo + (z_InLo + kDeltaZLum) * (rtRatio - 1.f) * (z_InLo < 0.f ? 1.f : dzDrtScale) + (zpIn + zpOut);
    const float zLo = z_InLo + (z_InLo - kDeltaZLum) * (rtRatio - 1.f) * (z_InLo > 0.f ? 1.f : dzDrtScale) - (zpIn + zpOut);
    if (tk) {  // Tracklet.cc:992-994 and the end's 1479-1488
        tk[TK_z_OutLo] = z_OutLo; tk[TK_zLo] = zLo; tk[TK_zHi] = zHi; tk[TK_z_InLo] = z_InLo;
        tk[TK_rt_InLo] = rt_InLo; tk[TK_rt_OutLo] = rt_OutLo; tk[TK_rtRatio_OutLoInLo] = rtRatio;
        tk[TK_alpha1GeV_OutLo] = L.alpha1; tk[TK_dzDrtScale] = dzDrtScale;
        tk[TK_zpitch_InLo] = zpIn; tk[TK_zpitch_OutLo] = zpOut; tk[TK_drt_InSeg] = I.d;
    }
    if (!(z_OutLo >= zLo && z_OutLo <= zHi)) return false;
    const float drt_OutLo_InLo = rt_OutLo - rt_InLo;
    const float drt_InSeg = I.d;
    const float dz_InSeg = hInUp.z - hInLo.z;
    const float dr3_InSeg = hInUp.r3 - hInLo.r3;
    const float coshEta = dr3_InSeg / drt_InSeg;
    float dzErr = (zpIn + zpOut) * (zpIn + zpOut) * 2.f;
    const float thetaMuls = float(double(0.015f) * sqrt(double(0.1f) + 0.2 * double(rt_OutLo - rt_InLo) / double(50.f)) *
                                  I.sqrtR3Rt);
    const float sdlMuls = thetaMuls * 3.f / kPtCut * 4.f;
    dzErr += sdlMuls * sdlMuls * drt_OutLo_InLo * drt_OutLo_InLo / 3.f * coshEta * coshEta;
    dzErr = sqrtF(dzErr);
    const float dzMean = dz_InSeg / drt_InSeg * drt_OutLo_InLo;
    const float zWindow = dzErr / drt_InSeg * drt_OutLo_InLo + (zpIn + zpOut);
    const float zLoP = z_InLo + dzMean * (z_InLo > 0.f ? 1.f : dzDrtScale) - zWindow;
    const float zHiP = z_InLo + dzMean * (z_InLo < 0.f ? 1.f : dzDrtScale) + zWindow;
    if (tk) { tk[TK_zLoPointed] = zLoP; tk[TK_zHiPointed] = zHiP; }  // :1066-1067
    if (!(z_OutLo >= zLoP && z_OutLo <= zHiP)) return false;
    const float sdlPVoff = 0.1f / rt_OutLo;
    const float sdlCut = float(double(L.alpha1) + sqrt(double(sdlMuls * sdlMuls + sdlPVoff * sdlPVoff)));
    const float dPhiMid = phiMpiPi(phiOf(hOutLo.x - hInLo.x, hOutLo.y - hInLo.y) - midPhi(hInLo, hOutLo));
    if (tk) { tk[TK_sdlCut] = sdlCut; tk[TK_dPhi] = dPhiMid; }  // :1099, 1152
    if (!(fabsF(dPhiMid) <= sdlCut)) return false;
    c.sdlCut = sdlCut;
    c.thetaMuls = thetaMuls;
    c.sinDPhiMid = sinF(dPhiMid);
    if (m) {
        m->w1 = (z_OutLo - zLo) / (zHi - zLo);
        m->w2 = (z_OutLo - zLoP) / (zHiP - zLoP);
        m->phi = dPhiMid / sdlCut;
    }
    return true;
}

// BBEE / EEEE: Tracklet.cc:1540-1771 and 2018-2270 (rt windows, pointing, midpoint deltaPhi)
template <bool EEEE>
__device__ __forceinline__ bool tlStage1EC(const TLInner& I, const TLOutLo& L, TLCarry& c, TLMargins* m,
                                           float* tk = nullptr) {
    const HitView& hInLo = I.lo;
    const HitView& hInUp = I.up;
    const HitView& hOutLo = L.a;
    const bool psInLo = I.psLo, psOutLo = L.ps, psInUp = I.psUp;
    const float rtIn = hInLo.rt, rtOut = hOutLo.rt, zIn = hInLo.z, zOut = hOutLo.z;
    const float dzDrtScale = L.dzScale;
    if (!(zIn * zOut > 0)) return false;
    const float dLum = copysignf(kDeltaZLum, zIn);
    float rtLo, rtHi;
    if (!EEEE) {
        const float zGeom = (psInLo ? kPixelPSZpitch : kStrip2SZpitch) + (psOutLo ? kPixelPSZpitch : kStrip2SZpitch);
        const float rtGeom1 = psOutLo ? kPixelPSZpitch : kStrip2SZpitch;
        const float zGeom1 = copysignf(zGeom, zIn);
        rtLo = rtIn * (1.f + (zOut - zIn - zGeom1) / (zIn + zGeom1 + dLum) / dzDrtScale) - rtGeom1;
        if (!(rtOut >= rtLo)) return false;
        float zInForHi = zIn - zGeom1 - dLum;
        if (zInForHi * zIn < 0) zInForHi = copysignf(0.1f, zIn);
        rtHi = rtIn * (1.f + (zOut - zIn + zGeom1) / zInForHi) + rtGeom1;
    } else {
        if (tk) {  // the end's records (Tracklet.cc:2602-2612), with the barrel z window it never cuts on (:2127)
            const float rtOut_o_rtIn = rtOut / rtIn;
            const float zGeom = (psInLo ? kPixelPSZpitch : kStrip2SZpitch) + (psOutLo ? kPixelPSZpitch : kStrip2SZpitch);
            tk[TK_rtIn] = rtIn; tk[TK_rtOut] = rtOut; tk[TK_rtOut_o_rtIn] = rtOut_o_rtIn; tk[TK_zIn] = zIn;
            tk[TK_zOut] = zOut; tk[TK_zGeom] = zGeom; tk[TK_sdlSlope] = L.alpha1; tk[TK_dzDrtScale] = dzDrtScale;
            tk[TK_sdIn_d] = I.d;
            tk[TK_zLo] = zIn + (zIn - kDeltaZLum) * (rtOut_o_rtIn - 1.f) * (zIn > 0.f ? 1.f : dzDrtScale) - zGeom;
        }
        const float rtGeom = (psInLo && psOutLo) ? 2.f * kPixelPSZpitch
                             : (psInLo || psOutLo) ? (kPixelPSZpitch + kStrip2SZpitch) : 2.f * kStrip2SZpitch;
        const float dz = zOut - zIn;
        rtLo = rtIn * (1.f + dz / (zIn + dLum) / dzDrtScale) - rtGeom;
        if (!(rtOut >= rtLo)) return false;
        rtHi = rtIn * (1.f + dz / (zIn - dLum)) + rtGeom;
    }
    if (!(rtOut >= rtLo && rtOut <= rtHi)) return false;
    const float drtSDIn = I.d;
    const float dzSDIn = hInUp.z - hInLo.z;
    const float dr3SDIn = hInUp.r3 - hInLo.r3;
    const float coshEta = dr3SDIn / drtSDIn;
    const float dzOutInAbs = fabsF(zOut - zIn);
    const float kZ = (zOut - zIn) / dzSDIn;
    float thetaMuls;
    if (!EEEE) {
        thetaMuls = float(double(0.015f) * sqrt(double(0.1f) + 0.2 * double(rtOut - rtIn) / double(50.f)) *
                          I.sqrtR3Rt);
        if (!(kZ >= 0 && rtOut >= rtLo && rtOut <= rtHi)) return false;
    } else {
        thetaMuls = float(double(0.015f) * sqrt(double(0.1f) + 0.2 * double(rtOut - rtIn) / double(50.f)));
    }
    const float sdlMuls = thetaMuls * 3.f / kPtCut * 4.f;
    float w2 = 0.5f;
    if (EEEE && psInLo && psInUp) {
        const float multDzDr = dzOutInAbs * coshEta / (coshEta * coshEta - 1.f);
        const float rtGeom = (psInLo && psOutLo) ? 2.f * kPixelPSZpitch
                             : (psInLo || psOutLo) ? (kPixelPSZpitch + kStrip2SZpitch) : 2.f * kStrip2SZpitch;
        float drtErr = kPixelPSZpitch * kPixelPSZpitch * 2.f / dzSDIn / dzSDIn * dzOutInAbs * dzOutInAbs;
        drtErr += sdlMuls * sdlMuls * multDzDr * multDzDr / 3.f * coshEta * coshEta;
        drtErr = sqrtF(drtErr);
        const float drtMean = drtSDIn * dzOutInAbs / fabsF(dzSDIn);
        const float rtWindow = drtErr + rtGeom;
        const float rtLoP = rtIn + drtMean / dzDrtScale - rtWindow;
        const float rtHiP = rtIn + drtMean + rtWindow;
        if (!(kZ >= 0 && rtOut >= rtLoP && rtOut <= rtHiP)) return false;
        if (m) w2 = (rtOut - rtLoP) / (rtHiP - rtLoP);
    }
    const float sdlPVoff = 0.1f / rtOut;
    const float sdlCut = float(double(L.alpha1) + sqrt(double(sdlMuls * sdlMuls + sdlPVoff * sdlPVoff)));
    const float dPhiMid = phiMpiPi(phiOf(hOutLo.x - hInLo.x, hOutLo.y - hInLo.y) - midPhi(hInLo, hOutLo));
    if (!(fabsF(dPhiMid) <= sdlCut)) return false;
    c.sdlCut = sdlCut;
    c.thetaMuls = thetaMuls;
    c.sinDPhiMid = EEEE ? 0.f : sinF(dPhiMid);  // the EEEE second stage has no dBetaROut term (Tracklet.cc:2493)
    if (m) {
        m->w1 = (rtOut - rtLo) / (rtHi - rtLo);
        m->w2 = w2;
        m->phi = dPhiMid / sdlCut;
    }
    return true;
}

__device__ __forceinline__ bool tlStage1(int variant, const TLInner& I, const TLOutLo& L, TLCarry& c,
                                         TLMargins* m = nullptr, float* tk = nullptr) {
    if (variant == TL_BBBB) return tlStage1BBBB(I, L, c, m, tk);
    if (variant == TL_BBEE) return tlStage1EC<false>(I, L, c, m, nullptr);  // BBEE records the iterations only
    if (variant == TL_EEEE) return tlStage1EC<true>(I, L, c, m, tk);
    return false;
}

// ---- second stage: cuts that involve the outer segment's outer MD
// _v2: Tracklet.cc:1097-1100 and 1156-1501
// rec (a passing TL's per-object values, kTLVars floats): getBetaIn, getBetaInCut,
// getBetaOut, getBetaOutCut, getDeltaBeta, getDeltaBetaCut, then the map keys.
// nm1: DefaultNm1_TLAlgo (Tracklet.cc:127-131, 1513-1536), the last cut (dBeta) waived.
__device__ __forceinline__ bool tlStage2BBBB(const TLInner& I, const TLOutLo& L, const TLOutUp& U, const TLCarry& c,
                                             float* rec = nullptr, bool nm1 = false) {
    const HitView& hInLo = I.lo;
    const HitView& hInUp = I.up;
    const HitView& hOutLo = L.a;
    const HitView& hOutUp = U.a;
    const float k2 = k2Rinv();
    float* tk = rec ? rec + kTLTyped : nullptr;
    if (tk) tk[TK_deltaPhiPos] = dPhi(hInUp.phi, hOutUp.phi);  // Tracklet.cc:1098
    if (!(fabsF(dPhi(hInUp.phi, hOutUp.phi)) <= c.sdlCut)) return false;
    const float alpha_InLo = I.dphichange;
    const bool isEC_lastLayer = U.ec2s;
    const float axX = hOutUp.x - hInLo.x, axY = hOutUp.y - hInLo.y;
    const float axPhi = phiOf(axX, axY);
    float betaIn = alpha_InLo - phiMpiPi(axPhi - hInLo.phi);
    float betaInRHmin = betaIn, betaInRHmax = betaIn;
    float betaOut = -U.aOU + phiMpiPi(axPhi - hOutUp.phi);
    float betaOutRHmin = betaOut, betaOutRHmax = betaOut;
    float hiRt = hOutUp.rt, loRt = hOutUp.rt;
    if (isEC_lastLayer) {
        betaOutRHmin = -U.aOUhi + phiMpiPi(phiOf(U.hiX - hInLo.x, U.hiY - hInLo.y) - U.hiPhi);
        betaOutRHmax = -U.aOUlo + phiMpiPi(phiOf(U.loX - hInLo.x, U.loY - hInLo.y) - U.loPhi);
        hiRt = U.hiRt;
        loRt = U.loRt;
    }
    const float drt_tl_axis = rtOf(axX, axY);
    const float rt_InSeg = I.dr;
    const float drt_InSeg = I.d;
    const float betaIn_cut = asinF(fmin2((-rt_InSeg * 1.f + drt_tl_axis) * k2 / kPtCut, kSinAlphaMax)) + (0.02f / drt_InSeg);
    if (!(fabsF(betaInRHmin) < betaIn_cut)) return false;
    if (rec) { rec[0] = betaInRHmin; rec[1] = betaIn_cut; }  // Tracklet.cc:1218-1219
    if (tk) {  // :1197-1204, 1217 (the RH ranges are recorded again, rescaled, at the end)
        tk[TK_betaIn_0th] = betaIn; tk[TK_betaOut_0th] = betaOut;
        tk[TK_rawBetaIn] = betaIn; tk[TK_rawBetaOut] = betaOut; tk[TK_betaIn_cut] = betaIn_cut;
    }
    float betaAv = 0.5f * (betaIn + betaOut);
    float pt_beta = crm::divSinCR(drt_tl_axis * k2, betaAv);
    const int lOut = isEC_lastLayer ? 11 : 5;
    const float sdOut_dr = U.dr;
    const float sdOut_d = U.d;
    deltaBetaIter(betaIn, betaOut, betaAv, pt_beta, rt_InSeg, sdOut_dr, drt_tl_axis, 5, tk);
    rescaleRH(betaIn, betaInRHmin, betaInRHmax);
    rescaleRH(betaOut, betaOutRHmin, betaOutRHmax);
    const float dBetaMuls = c.thetaMuls * 4.f / fmin2(fabsF(pt_beta), 7.0f);
    const float alphaOutAbsReg = fmax2(fabsF(U.dphichange), L.alphaReg3);
    const float dBetaOutLum = lOut < 11 ? 0.0f : fabsF(alphaOutAbsReg * kDeltaZLum / hOutLo.z);
    const float dBetaLum2 = (0.0f + dBetaOutLum) * (0.0f + dBetaOutLum);
    const float sinDPhi = c.sinDPhiMid;
    const float dBetaROut2 = float(sq(double((hiRt - loRt) * sinDPhi / drt_tl_axis)));
    const float betaOut_cut = float(double(asinF(fmin2(drt_tl_axis * k2 / kPtCut, kSinAlphaMax)) + (0.02f / sdOut_d)) +
                                    sqrt(double(dBetaLum2 + dBetaMuls * dBetaMuls)));
    if (!(fabsF(betaOut) < betaOut_cut)) return false;
    if (rec) { rec[2] = betaOut; rec[3] = betaOut_cut; }  // Tracklet.cc:1375-1376
    const float dBetaRes = 0.02f / fmin2(sdOut_d, drt_InSeg);
    const float dBetaCut2 = float(double(dBetaRes * dBetaRes * 2.0f + dBetaMuls * dBetaMuls + dBetaLum2 + 0.f + dBetaROut2) +
                                  0.25 * sq(double(fabsF(betaInRHmin - betaInRHmax) + fabsF(betaOutRHmin - betaOutRHmax))));
    const float dBeta = betaIn - betaOut;
    if (rec) { rec[4] = dBeta; rec[5] = sqrtF(dBetaCut2); }  // Tracklet.cc:1498-1499
    if (tk) {  // :1438-1493
        tk[TK_betaIn] = betaIn; tk[TK_betaInRHmin] = betaInRHmin; tk[TK_betaInRHmax] = betaInRHmax;
        tk[TK_betaOut] = betaOut; tk[TK_betaOutRHmin] = betaOutRHmin; tk[TK_betaOutRHmax] = betaOutRHmax;
        tk[TK_dBeta] = dBeta; tk[TK_dBetaCut2] = dBetaCut2; tk[TK_dBetaLum2] = dBetaLum2;
        tk[TK_dBetaMuls] = dBetaMuls; tk[TK_dBetaRIn2] = 0.f; tk[TK_dBetaROut2] = dBetaROut2;
        tk[TK_dBetaRes] = dBetaRes; tk[TK_drt_tl_axis] = drt_tl_axis; tk[TK_sdOut_d] = sdOut_d;
        tk[TK_pt_betaIn] = crm::divSinCR(drt_tl_axis * k2, betaIn);  // :1393-1394
        tk[TK_pt_betaOut] = crm::divSinCR(drt_tl_axis * k2, betaOut);
        tk[TK_pt_beta] = pt_beta; tk[TK_betaAv] = betaAv;
    }
    return nm1 || dBeta * dBeta <= dBetaCut2;
}

// BBEE: Tracklet.cc:1774-2016; EEEE: Tracklet.cc:2273-2665
template <bool EEEE>
__device__ __forceinline__ bool tlStage2EC(const TLInner& I, const TLOutLo& L, const TLOutUp& U, const TLCarry& c,
                                           float* rec = nullptr, bool nm1 = false) {
    const HitView& hInLo = I.lo;
    const HitView& hInUp = I.up;
    const HitView& hOutLo = L.a;
    const HitView& hOutUp = U.a;
    const float k2 = k2Rinv();
    float* tk = rec ? rec + kTLTyped : nullptr;
    if (!(fabsF(dPhi(hInUp.phi, hOutUp.phi)) <= c.sdlCut)) return false;
    const float sdIn_alpha = I.dphichange;
    const float sdOut_alpha = I.dphichange;  // A-8
    const float dr3x = hOutUp.x - hInLo.x, dr3y = hOutUp.y - hInLo.y;
    const float dr3Phi = phiOf(dr3x, dr3y);
    float betaIn = sdIn_alpha - phiMpiPi(dr3Phi - hInLo.phi);
    float betaInRHmin, betaInRHmax, betaOutRHmin, betaOutRHmax, betaOut;
    if (!EEEE) {
        const float sdOut_alphaOut = U.aOU;
        const float aMin = phiMpiPi(U.dphichangemin - U.dphimin);
        const float aMax = phiMpiPi(U.dphichangemax - U.dphimax);
        betaOut = -sdOut_alphaOut + phiMpiPi(dr3Phi - hOutUp.phi);
        betaInRHmin = betaIn;  // the 2S test never holds (A-9)
        betaInRHmax = betaIn;
        betaOutRHmin = betaOut - aMin + sdOut_alphaOut;
        betaOutRHmax = betaOut - aMax + sdOut_alphaOut;
        if (fabsF(betaOutRHmin) > fabsF(betaOutRHmax)) { float t = betaOutRHmin; betaOutRHmin = betaOutRHmax; betaOutRHmax = t; }
    } else {
        const float sdOut_dPhiPos = U.dPhiPos;
        const float aMin = phiMpiPi(U.dphichangemin - sdOut_dPhiPos);
        const float aMax = phiMpiPi(U.dphichangemax - sdOut_dPhiPos);
        const float sdOut_alphaOut = phiMpiPi(U.dphichange - sdOut_dPhiPos);
        betaOut = -sdOut_alphaOut + phiMpiPi(dr3Phi - hOutUp.phi);
        if (tk) {  // Tracklet.cc:2303-2308
            tk[TK_sdIn_alpha] = sdIn_alpha; tk[TK_sdOut_alphaOut] = sdOut_alphaOut;
            tk[TK_rawBetaInCorrection] = phiMpiPi(dr3Phi - hInLo.phi);
            tk[TK_rawBetaOutCorrection] = phiMpiPi(dr3Phi - hOutUp.phi);
            tk[TK_rawBetaIn] = betaIn; tk[TK_rawBetaOut] = betaOut;
        }
        betaInRHmin = betaIn + I.dphichangemin - sdIn_alpha;
        betaInRHmax = betaIn + I.dphichangemax - sdIn_alpha;
        if (fabsF(betaInRHmin) > fabsF(betaInRHmax)) { float t = betaInRHmin; betaInRHmin = betaInRHmax; betaInRHmax = t; }
        betaOutRHmin = betaOut - aMin + sdOut_alphaOut;
        betaOutRHmax = betaOut - aMax + sdOut_alphaOut;
        if (fabsF(betaOutRHmin) > fabsF(betaOutRHmax)) { float t = betaOutRHmin; betaOutRHmin = betaOutRHmax; betaOutRHmax = t; }
    }
    const float sdIn_dr = I.dr;
    const float sdIn_d = I.d;
    const float dr = rtOf(dr3x, dr3y);
    if (!EEEE) {
        const float betaIn_cut = asinF(fmin2((-sdIn_dr * 1.f + dr) * k2 / kPtCut, kSinAlphaMax)) + (0.02f / sdIn_d);
        if (!(fabsF(betaInRHmin) < betaIn_cut)) return false;
    }  // in EEEE the betaIn cut never rejects (A-11)
    if (EEEE && rec) {  // its value has no asin (A-12; Tracklet.cc:2319-2324)
        rec[0] = betaInRHmin;
        rec[1] = (-sdIn_dr * 1.f + dr) * k2 / kPtCut + (0.02f / sdIn_d);
    }
    float betaAv = 0.5f * (betaIn + betaOut);
    float pt_beta = crm::divSinCR(dr * k2, betaAv);
    const int lIn = EEEE ? 11 : 5;
    const float sdOut_dr = U.dr;
    const float sdOut_d = U.d;
    deltaBetaIter(betaIn, betaOut, betaAv, pt_beta, sdIn_dr, sdOut_dr, dr, lIn, tk);  // BBEE's only map keys
    rescaleRH(betaIn, betaInRHmin, betaInRHmax);
    rescaleRH(betaOut, betaOutRHmin, betaOutRHmax);
    const float dBetaMuls = c.thetaMuls * 4.f / fmin2(fabsF(pt_beta), 7.0f);
    const float alphaOutAbsReg = fmax2(fabsF(sdOut_alpha), L.alphaReg3);
    float dBetaInLum = 0.0f;
    if (EEEE) {
        const float alphaInAbsReg = fmax2(fabsF(sdIn_alpha), I.alphaReg3Lo);
        dBetaInLum = fabsF(alphaInAbsReg * kDeltaZLum / hInLo.z);
    }
    const float dBetaOutLum = fabsF(alphaOutAbsReg * kDeltaZLum / hOutLo.z);
    const float dBetaLum2 = (dBetaInLum + dBetaOutLum) * (dBetaInLum + dBetaOutLum);
    float dBetaROut2 = 0.f;
    if (!EEEE) {
        const float sinDPhi = c.sinDPhiMid;
        const float hiRt = U.ec2s ? U.hiRt : hOutUp.rt;
        const float loRt = U.ec2s ? U.loRt : hOutUp.rt;
        dBetaROut2 = float(sq(double((hiRt - loRt) * sinDPhi / dr)));
    }
    const float betaOut_cut = float(double(asinF(fmin2(dr * k2 / kPtCut, kSinAlphaMax)) + (0.02f / sdOut_d)) +
                                    sqrt(double(dBetaLum2 + dBetaMuls * dBetaMuls)));
    if (!(fabsF(betaOut) < betaOut_cut)) return false;
    if (EEEE && rec) { rec[2] = betaOut; rec[3] = betaOut_cut; }  // Tracklet.cc:2499-2500
    const float dBetaRes = 0.02f / fmin2(sdOut_d, sdIn_d);
    const float dBetaCut2 = float(double(dBetaRes * dBetaRes * 2.0f + dBetaMuls * dBetaMuls + dBetaLum2 + 0.f + dBetaROut2) +
                                  0.25 * sq(double(fabsF(betaInRHmin - betaInRHmax) + fabsF(betaOutRHmin - betaOutRHmax))));
    const float dBeta = betaIn - betaOut;
    if (rec) {  // BBEE sets only deltaBeta (Tracklet.cc:1977); EEEE both (2621-2622)
        rec[4] = dBeta;
        if (EEEE) rec[5] = sqrtF(dBetaCut2);
    }
    if (EEEE && tk) {  // :2576-2617
        tk[TK_betaIn] = betaIn; tk[TK_betaInRHmin] = betaInRHmin; tk[TK_betaInRHmax] = betaInRHmax;
        tk[TK_betaOut] = betaOut; tk[TK_betaOutRHmin] = betaOutRHmin; tk[TK_betaOutRHmax] = betaOutRHmax;
        tk[TK_dBeta] = dBeta; tk[TK_dBetaCut2] = dBetaCut2; tk[TK_dBetaLum2] = dBetaLum2;
        tk[TK_dBetaMuls] = dBetaMuls; tk[TK_dBetaRIn2] = 0.f; tk[TK_dBetaROut2] = 0.f;
        tk[TK_dBetaRes] = dBetaRes; tk[TK_dr] = dr; tk[TK_sdOut_d] = sdOut_d;
        tk[TK_pt_betaIn] = crm::divSinCR(dr * k2, betaIn);  // :2516-2517
        tk[TK_pt_betaOut] = crm::divSinCR(dr * k2, betaOut);
        tk[TK_pt_beta] = pt_beta; tk[TK_betaAv] = betaAv;
    }
    return nm1 || dBeta * dBeta <= dBetaCut2;
}

// bOutUp: the outer segment's outer MD is in the barrel (the fourth subdet of the dispatch)
__device__ __forceinline__ bool tlStage2(int variant, bool bOutUp, const TLInner& I, const TLOutLo& L,
                                         const TLOutUp& U, const TLCarry& c, float* rec = nullptr,
                                         bool nm1 = false) {
    if (variant == TL_BBBB) return tlStage2BBBB(I, L, U, c, rec, nm1);
    if (bOutUp) return false;
    if (variant == TL_BBEE) return tlStage2EC<false>(I, L, U, c, rec, nm1);
    return tlStage2EC<true>(I, L, U, c, rec, nm1);
}

}  // namespace sdl
