// CPU ORACLE for the SDL segment-linking chain -- test infrastructure only.
// See sdl_oracle.h for its role and the two libm policies.  Every function
// below cites the reference file:line whose behaviour it restates.
#include "sdl_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace oracle {

constexpr double kPI = 3.14159265358979323846;  // M_PI
constexpr float kSlopeInfF = 123456789.0f;       // SDL_INF compared as float (MiniDoublet.cc:3,1130)

// float(2.99792458e-3 * 3.8) and its exact half (MiniDoublet.cc:746-747)
const float kRinv1GeVf = float(2.99792458e-3 * 3.8);
const float k2Rinv1GeVf = float(double(kRinv1GeVf) / 2.);
constexpr float kPtCut = 1.0f;        // PTCUT (Constants.h:4)
constexpr float kSinAlphaMax = 0.95f;
constexpr float kDeltaZLum = 15.f;
constexpr float kPixelPSZpitch = 0.15f;
constexpr float kStrip2SZpitch = 5.0f;

// ------------------------------------------------------------------ libm policy
struct Libm {
    int mode;  // ORACLE_LIBM_*
    float pick(float glibc, double exact) const {
        if (mode == ORACLE_LIBM_GLIBC) return glibc;
        const float r = float(exact);
        if (mode == ORACLE_LIBM_CR) return r;
        return r * (1.f + 1.f / 256.f);  // planted error (tests): 0.4 % too large
    }
    float asinF(float v) const { return mode == ORACLE_LIBM_GLIBC ? std::asin(v) : pick(0.f, std::asin(double(v))); }
    float atanF(float v) const { return mode == ORACLE_LIBM_GLIBC ? std::atan(v) : pick(0.f, std::atan(double(v))); }
    float sinF(float v) const { return mode == ORACLE_LIBM_GLIBC ? std::sin(v) : pick(0.f, std::sin(double(v))); }
    float cosF(float v) const { return mode == ORACLE_LIBM_GLIBC ? std::cos(v) : pick(0.f, std::cos(double(v))); }
    float tanF(float v) const { return mode == ORACLE_LIBM_GLIBC ? std::tan(v) : pick(0.f, std::tan(double(v))); }
    float acosF(float v) const { return mode == ORACLE_LIBM_GLIBC ? std::acos(v) : pick(0.f, std::acos(double(v))); }
};

// Keys of a tracklet's reco-variable map (Tracklet.cc:152-248 resets 97 of them
// to -999, :256-293 sets "algo", _v2 adds ten of its own), in std::map order:
// column kTLTyped + key of the per-object values.
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
static_assert(int(kTLKeyCount) == int(kTLKeys), "tracklet key count");
// the keys only runTrackletDefaultAlgoBarrelBarrelBarrelBarrel_v2 writes (absent elsewhere)
const int kTLv2Keys[] = {TK_alpha1GeV_OutLo, TK_drt_InSeg, TK_drt_tl_axis, TK_rtRatio_OutLoInLo, TK_rt_InLo,
                         TK_rt_OutLo, TK_z_InLo, TK_z_OutLo, TK_zpitch_InLo, TK_zpitch_OutLo};

// std::min / std::max semantics (NaN handling follows the argument order).
inline float fmin2(float a, float b) { return (b < a) ? b : a; }
inline float fmax2(float a, float b) { return (a < b) ? b : a; }
// pow(x, 2) on a float argument: exact in double.
inline double sq(double v) { return v * v; }

// Unqualified abs(float) binds to int abs(int) (A-1): truncation toward zero,
// with the x86 cvttss2si result INT_MIN for NaN / out-of-range values.
inline int absTrunc(float v) {
    int i = (v > -2147483648.f && v < 2147483648.f) ? int(v) : INT_MIN;
    return i == INT_MIN ? INT_MIN : (i < 0 ? -i : i);
}

// MathUtil.cc:3-19
inline float phiMpiPi(float x) {
    if (std::isnan(x) || std::isinf(x)) return x;
    while (double(x) >= kPI) x = float(double(x) - 2. * kPI);
    while (double(x) < -kPI) x = float(double(x) + 2. * kPI);
    return x;
}
// MathUtil.cc:21-27
inline float atan2Sdl(float y, float x) {
    if (x != 0) return float(std::atan2(double(y), double(x)));
    if (y == 0) return 0.f;
    if (y > 0) return float(kPI / 2);
    return float(-kPI / 2);
}

// A point with the derived quantities of Hit::setDerivedQuantities (Hit.cc:68-83).
struct P3 {
    float x, y, z;
    float rt() const { return float(std::sqrt(double(x * x + y * y))); }
    float r3() const { return float(std::sqrt(double(x * x + y * y + z * z))); }
    float phi() const { return phiMpiPi(float(kPI + double(atan2Sdl(-y, -x)))); }
};
inline P3 sub(const P3& a, const P3& b) { return P3{a.x - b.x, a.y - b.y, a.z - b.z}; }  // Hit.cc:195-198
// Hit() then += a, += b, /= 2 (Tracklet.cc:1131-1134)
inline P3 midpoint(const P3& a, const P3& b) {
    P3 m{0.f + a.x, 0.f + a.y, 0.f + a.z};
    m.x = (m.x + b.x) / 2.0f; m.y = (m.y + b.y) / 2.0f; m.z = (m.z + b.z) / 2.0f;
    return m;
}
// Hit::deltaPhi (Hit.cc:149-152): phi of the argument minus phi of this
inline float deltaPhi(float phiA, const P3& b) { return phiMpiPi(b.phi() - phiA); }

struct HitD {
    P3 p;
    float rt, r3, phi;
    int idx;       // user hit index
    int slot;      // module slot in the event
    int hiEdge = -1, loEdge = -1;  // indices into Event::edges
};

// ------------------------------------------------------------------ modules
enum { SUB_ENDCAP = 4, SUB_BARREL = 5 };
enum { SIDE_NEGZ = 1, SIDE_POSZ = 2, SIDE_CENTER = 3 };
enum { TYPE_PS = 0, TYPE_2S = 1 };
enum { LT_PIXEL = 0, LT_STRIP = 1 };

struct ModInfo {
    uint32_t detId, partner;
    int subdet, side, layer, rod, ring, module;
    bool isLower, isInverted;
    int type, layerType;
};

// Module.cc:159-358, restated.
ModInfo decode(uint32_t d) {
    ModInfo m{};
    m.detId = d;
    m.subdet = (d >> 25) & 7;
    if (m.subdet == SUB_ENDCAP) { m.side = (d >> 23) & 3; m.layer = (d >> 18) & 7; m.ring = (d >> 12) & 15; m.rod = 0; }
    else if (m.subdet == SUB_BARREL) { m.side = (d >> 18) & 3; m.layer = (d >> 20) & 7; m.rod = (d >> 10) & 127; m.ring = 0; }
    else { m.side = 0; m.layer = 0; m.rod = 0; m.ring = 0; }
    m.module = (d >> 2) & 127;
    bool inv = false;
    if (m.subdet == SUB_ENDCAP) {
        if (m.side == SIDE_NEGZ) inv = m.module % 2 == 1;
        else if (m.side == SIDE_POSZ) inv = m.module % 2 == 0;
    } else if (m.subdet == SUB_BARREL) {
        if (m.side == SIDE_CENTER) {
            if (m.layer <= 3) inv = m.module % 2 == 1;
            else if (m.layer >= 4) inv = m.module % 2 == 0;
        } else if (m.side == SIDE_NEGZ || m.side == SIDE_POSZ) {
            if (m.layer <= 2) inv = m.module % 2 == 1;
            else if (m.layer == 3) inv = m.module % 2 == 0;
        }
    }
    m.isInverted = inv;
    m.isLower = inv ? !(d & 1) : (d & 1);
    if (m.isLower) m.partner = inv ? d - 1 : d + 1;
    else m.partner = inv ? d + 1 : d - 1;
    if (m.subdet == SUB_BARREL) m.type = m.layer <= 3 ? TYPE_PS : TYPE_2S;
    else if (m.layer <= 2) m.type = m.ring <= 10 ? TYPE_PS : TYPE_2S;
    else m.type = m.ring <= 7 ? TYPE_PS : TYPE_2S;
    if (m.type == TYPE_2S) m.layerType = LT_STRIP;
    else if (inv) m.layerType = m.isLower ? LT_STRIP : LT_PIXEL;
    else m.layerType = m.isLower ? LT_PIXEL : LT_STRIP;
    return m;
}

// MiniDoublet.cc:1242-1257
bool isTighterTilted(const ModInfo& m) {
    if (m.subdet != SUB_BARREL || m.side == SIDE_CENTER) return false;
    if (m.layer == 3) return true;
    if (m.side == SIDE_NEGZ && m.layer == 2 && m.rod > 5) return true;
    if (m.side == SIDE_POSZ && m.layer == 2 && m.rod < 8) return true;
    if (m.side == SIDE_NEGZ && m.layer == 1 && m.rod > 9) return true;
    if (m.side == SIDE_POSZ && m.layer == 1 && m.rod < 4) return true;
    return false;
}
// MiniDoublet.cc:1260-1332
float moduleGapSize(const ModInfo& m) {
    static const float flat[6] = {0.26f, 0.16f, 0.16f, 0.18f, 0.18f, 0.18f};
    int iL = m.layer - 1;
    if (m.subdet == SUB_BARREL && m.side == SIDE_CENTER) return flat[iL];
    if (isTighterTilted(m)) return 0.26f;
    if (m.subdet == SUB_ENDCAP) {
        int j = m.ring - 1;
        if (iL == 0 || iL == 1) return j < 10 ? 0.4f : 0.18f;
        if (iL == 2 || iL == 3) return j < 8 ? 0.4f : 0.18f;
        return j < 9 ? 0.4f : 0.18f;
    }
    return 0.4f;
}

// ------------------------------------------------------------------ geometry
struct Geometry {
    std::unordered_map<uint32_t, float> drdz, slope, ecSlopeLower, ecCentroidPhi;
    std::unordered_map<uint32_t, std::vector<uint32_t>> conn;
    // A missing key reads as 0 / empty (std::map::operator[] in TiltedGeometry.cc:45,50,
    // EndcapGeometry.cc:65-100, ModuleConnectionMap.cc:105).
    static float get(const std::unordered_map<uint32_t, float>& m, uint32_t k) {
        auto it = m.find(k);
        return it == m.end() ? 0.f : it->second;
    }
    const std::vector<uint32_t>& connected(uint32_t k) const {
        static const std::vector<uint32_t> none;
        auto it = conn.find(k);
        return it == conn.end() ? none : it->second;
    }
};

bool loadGeometry(Geometry& g, const char* tilted, const char* endcap, const char* map) {
    std::ifstream ft(tilted), fe(endcap), fm(map);
    if (!ft || !fe || !fm) return false;
    std::string line;
    while (std::getline(ft, line)) {  // TiltedGeometry.cc:26-41
        std::stringstream ss(line);
        unsigned int d; float a, b;
        ss >> d >> a >> b;
        g.drdz[d] = a; g.slope[d] = b;
    }
    while (std::getline(fe, line)) {  // EndcapGeometry.cc:33-60 (column 7 is stored as phi)
        std::stringstream ss(line);
        unsigned int d; float avgr2, yl, sl, yh, sh, cr, cp, cz;
        ss >> d >> avgr2 >> yl >> sl >> yh >> sh >> cr >> cp >> cz;
        g.ecSlopeLower[d] = sl; g.ecCentroidPhi[d] = cr;
    }
    while (std::getline(fm, line)) {  // ModuleConnectionMap.cc:26-46
        std::stringstream ss(line);
        unsigned int d; int n;
        ss >> d >> n;
        std::vector<uint32_t> v;
        for (int i = 0; i < n; ++i) { unsigned int c; ss >> c; v.push_back(c); }
        g.conn[d] = v;
    }
    return true;
}

// ------------------------------------------------------------------ tracing
struct Trace {
    int n = 0;
    int failed = -1;
    double v[32], lo[32], hi[32];
    int id[32];
    void add(int cid, double var, double l, double h, bool pass) {
        if (n < 32) { id[n] = cid; v[n] = var; lo[n] = l; hi[n] = h; ++n; }
        if (!pass && failed < 0) failed = cid;
    }
};
const double kNone = std::nan("");
#define CUT(cid, cond, var, l, h)                        \
    do {                                                 \
        bool _pass = (cond);                             \
        if (tr) tr->add(cid, var, l, h, _pass);          \
        if (!_pass) return false;                        \
    } while (0)

// ------------------------------------------------------------------ objects
struct MDRec {
    int lo, up, anchor;  // hit indices in the event
    int slot;            // lower module slot
    float dz, shiftedDz, dphi, dphichange;
    // made without running cuts (createPseudoMiniDoubletsFromAnchorModule, or
    // the AllComb algorithm): every cached value 0, no reco variable
    bool pseudo = false;
};
struct SGRec {
    int mdIn, mdOut;
    float dphi, dphimin, dphimax, dphichange, dphichangemin, dphichangemax;
    bool allComb = false;  // accepted by AllComb_SGAlgo: constructor values, no reco variable
};
struct PairRec {
    int in, out;
    bool allComb = false;  // a TL or TC accepted by the AllComb algorithm (no value set)
};

struct Slot {
    ModInfo info;
    float drdzStrip = 0, slopeStrip = 0, ecSlopeStrip = 0, centroidPhi = 0, gap = 0;
    std::vector<int> hits, mds, sgs, tps, tls;
};

struct Event {
    const Geometry* g;
    Libm lm;
    std::vector<HitD> hits, edges;
    std::vector<Slot> slots;
    std::unordered_map<uint32_t, int> slotOf;
    std::vector<int> lowerOrder;
    std::vector<MDRec> mds;
    std::vector<SGRec> sgs;
    std::vector<PairRec> tps, tls;
    std::vector<std::pair<int, PairRec>> tcs;  // (0 = from triplets, 1 = from tracklets)
    unsigned long long nCand[5] = {0, 0, 0, 0, 0};
    // Per-layer counters (Event.cc:1193-1344): kinds 0-4 candidates, 5-9 stored
    // objects (MD, SG, TP, TL, TC); bins 0-5 barrel layers 1-6, 6-10 endcap
    // layers 1-5, of the module (or, in the layer-pair stages, the inner layer)
    // each create* loop passes to the reference's increment functions.
    unsigned long long layerCnt[10][11] = {};
    void countAt(int kind, int bin) {
        if (bin >= 0 && bin < 11) layerCnt[kind][bin]++;
    }
    int binOfSlot(int s) const {
        const ModInfo& m = slots[s].info;
        if (m.subdet == SUB_BARREL && m.layer >= 1 && m.layer <= 6) return m.layer - 1;
        if (m.subdet == SUB_ENDCAP && m.layer >= 1 && m.layer <= 5) return 6 + m.layer - 1;
        return -1;
    }
    // While a stored object is re-evaluated for its per-object variables
    // (objectVars), the algorithms write them here (layouts: kMDVars, kSGVars,
    // kTLVars); null otherwise.
    mutable float* rec = nullptr;
    // DefaultNm1_TLAlgo (Tracklet.cc:127-131): the default tracklet algorithm
    // with its last cut, dBeta, waived (Tracklet.cc:1513-1536, 1990-2014, 2636-2653)
    bool tlNm1 = false;
    // the stage being run uses the AllComb algorithm (accepts every candidate)
    bool acMD = false, acSG = false, acTP = false, acTL = false, acTC = false;
    // Segment::runSegmentAllCombAlgo (Segment.cc:313-335): the pass bit only; the
    // segment keeps its constructor values (Segment.cc:35-56)
    bool allCombSG(int mi, int mo, SGRec& sg) const {
        sg.mdIn = mi;
        sg.mdOut = mo;
        sg.dphi = sg.dphimin = sg.dphimax = sg.dphichange = sg.dphichangemin = sg.dphichangemax = 0;
        sg.allComb = true;
        return true;
    }

    // Event::getModule (Event.cc:54-78)
    int getSlot(uint32_t d) {
        auto it = slotOf.find(d);
        if (it != slotOf.end()) return it->second;
        int s = int(slots.size());
        Slot sl;
        sl.info = decode(d);
        uint32_t stripId = sl.info.layerType == LT_PIXEL ? sl.info.partner : d;
        sl.drdzStrip = Geometry::get(g->drdz, stripId);
        sl.slopeStrip = Geometry::get(g->slope, stripId);
        sl.ecSlopeStrip = Geometry::get(g->ecSlopeLower, stripId);
        sl.centroidPhi = Geometry::get(g->ecCentroidPhi, d);
        sl.gap = moduleGapSize(sl.info);
        slots.push_back(sl);
        slotOf[d] = s;
        if (slots[s].info.isLower) lowerOrder.push_back(s);
        return s;
    }
    int findSlot(uint32_t d) const {
        auto it = slotOf.find(d);
        return it == slotOf.end() ? -1 : it->second;
    }

    // GeometryUtil.cc:3-23: hit + sign * 2.5 * (cos phi_c, sin phi_c, 0)
    HitD edgeHit(const HitD& h, bool up) const {
        float phic = slots[h.slot].centroidPhi;
        float sign = up ? 1.f : -1.f;
        P3 e{float(double(sign) * 2.5 * std::cos(double(phic))), float(double(sign) * 2.5 * std::sin(double(phic))), 0.f};
        e.x += h.p.x; e.y += h.p.y; e.z += h.p.z;
        HitD r;
        r.p = e; r.rt = e.rt(); r.r3 = e.r3(); r.phi = e.phi(); r.idx = -1; r.slot = h.slot;
        return r;
    }

    // Event::addHitToModule (Event.cc:120-139)
    void addHit(float x, float y, float z, int idx, uint32_t d) {
        HitD h;
        h.p = P3{x, y, z};
        h.rt = h.p.rt(); h.r3 = h.p.r3(); h.phi = h.p.phi();
        h.idx = idx;
        h.slot = getSlot(d);
        int hi = int(hits.size());
        hits.push_back(h);
        slots[h.slot].hits.push_back(hi);
        const ModInfo& m = slots[h.slot].info;
        if (m.subdet == SUB_ENDCAP && m.type == TYPE_2S) {
            edges.push_back(edgeHit(hits[hi], true));
            hits[hi].hiEdge = int(edges.size()) - 1;
            edges.push_back(edgeHit(hits[hi], false));
            hits[hi].loEdge = int(edges.size()) - 1;
        }
    }

    // ---------------------------------------------------------------- MD
    // MiniDoublet::dPhiThreshold (MiniDoublet.cc:741-810)
    float dPhiThreshold(const HitD& h, const Slot& s, float dPhi = 0, float dz = 1) const {
        static const float mulsB[6] = {0.0052f, 0.0038f, 0.0034f, 0.0034f, 0.0032f, 0.0034f};
        static const float mulsE[5] = {0.006f, 0.006f, 0.006f, 0.006f, 0.006f};
        static const float rB[6] = {21.8f, 34.6f, 49.6f, 67.4f, 87.6f, 106.8f};
        static const float rE[5] = {131.4f, 156.2f, 185.6f, 220.3f, 261.5f};
        const ModInfo& m = s.info;
        const float ptCut = 1;
        float rt = h.rt;
        unsigned iL = m.layer - 1;
        const float miniSlope = lm.asinF(fmin2(rt * k2Rinv1GeVf / ptCut, kSinAlphaMax));
        const float rLay = m.subdet == SUB_BARREL ? rB[iL] : rE[iL];
        const float miniPVoff = float(0.1 / double(rLay));
        const float miniMuls = m.subdet == SUB_BARREL ? mulsB[iL] * 3.f / ptCut : mulsE[iL] * 3.f / ptCut;
        const bool isTilted = m.subdet == SUB_BARREL && m.side != SIDE_CENTER;
        const float drdz = s.drdzStrip;
        const float miniTilt =
            isTilted ? float(double(0.5f * kPixelPSZpitch * drdz) / std::sqrt(double(1.f + drdz * drdz)) / double(s.gap))
                     : 0.f;
        const float miniLum = float(absTrunc(dPhi * kDeltaZLum / dz));
        if (m.subdet == SUB_BARREL && m.side == SIDE_CENTER)
            return float(double(miniSlope) + std::sqrt(sq(miniMuls) + sq(miniPVoff)));
        if (m.subdet == SUB_BARREL)
            return float(double(miniSlope) + std::sqrt(sq(miniMuls) + sq(miniPVoff) + sq(miniTilt * miniSlope)));
        return float(double(miniSlope) + std::sqrt(sq(miniMuls) + sq(miniPVoff) + sq(miniLum)));
    }

    // MiniDoublet::shiftStripHits (MiniDoublet.cc:989-1209)
    void shiftStripHits(const HitD& lower, const HitD& upper, const Slot& s, float& xn, float& yn, float& zn) const {
        const ModInfo& m = s.info;
        const HitD *pix, *str;
        if (m.type == TYPE_PS) {
            if (m.layerType == LT_PIXEL) { pix = &lower; str = &upper; }
            else { pix = &upper; str = &lower; }
        } else { pix = &lower; str = &upper; }
        bool isEndcap = m.subdet == SUB_ENDCAP;
        float angleA = float(std::fabs(double(lm.atanF(pix->rt / pix->p.z))));
        float angleB = isEndcap ? float(kPI / 2.) : lm.atanF(s.drdzStrip);
        float sep = s.gap;
        if (m.type == TYPE_PS && m.layerType != LT_PIXEL) sep *= -1;
        float drprime = (sep / lm.sinF(angleA + angleB)) * lm.sinF(angleA);
        float slope = isEndcap ? s.ecSlopeStrip : s.slopeStrip;
        float absArctanSlope = slope != kSlopeInfF ? float(std::fabs(double(lm.atanF(slope)))) : float(kPI / 2);
        float xp = pix->p.x, yp = pix->p.y;
        float angleM;
        if (xp > 0 && yp > 0) angleM = absArctanSlope;
        else if (xp > 0 && yp < 0) angleM = float(kPI - double(absArctanSlope));
        else if (xp < 0 && yp < 0) angleM = float(kPI + double(absArctanSlope));
        else angleM = float(2 * kPI - double(absArctanSlope));
        float dx = drprime * lm.sinF(angleM), dy = drprime * lm.cosF(angleM);
        float xa = xp + dx, ya = yp + dy;
        float xo = str->p.x, yo = str->p.y;
        if (slope == kSlopeInfF) { xn = xa; yn = yo; }
        else if (slope == 0) { xn = xo; yn = ya; }
        else {
            xn = (slope * xa + (1.f / slope) * xo - ya + yo) / (slope + (1.f / slope));
            yn = (xn - xa) * slope + ya;
        }
        float absdzprime = float(std::fabs(double(sep / lm.sinF(angleA + angleB) * lm.cosF(angleA))));
        float abszn = m.layerType == LT_PIXEL ? float(std::fabs(double(pix->p.z)) + double(absdzprime))
                                              : float(std::fabs(double(pix->p.z)) - double(absdzprime));
        zn = abszn * ((pix->p.z > 0) ? 1 : -1);
    }

    static float dPhiChange(const P3& a, float phiA, const P3& b) { return deltaPhi(phiA, sub(b, a)); }

    // MiniDoublet::runMiniDoubletDefaultAlgoBarrel (MiniDoublet.cc:243-447)
    bool mdBarrel(const HitD& lo, const HitD& up, const Slot& s, MDRec& md, Trace* tr) const {
        const ModInfo& m = s.info;
        md.dz = lo.p.z - up.p.z;
        const float dz = md.dz;
        const float dzCut = m.type == TYPE_PS ? 2.f : 10.f;
        const float sign = float(((dz > 0) - (dz < 0)) * ((lo.p.z > 0) - (lo.p.z < 0)));
        const float invertedcrossercut = float(absTrunc(dz) > 2) * sign;
        CUT(0, std::abs(dz) < dzCut && invertedcrossercut <= 0, std::abs(dz), kNone, dzCut);
        float miniCut = m.layerType == LT_PIXEL ? dPhiThreshold(lo, s) : dPhiThreshold(up, s);
        bool tilted = m.side != SIDE_CENTER;
        P3 shifted{0, 0, 0};
        float shiftedPhi = 0, shiftedRt = 0;
        if (tilted) {
            float xn, yn, zn;
            shiftStripHits(lo, up, s, xn, yn, zn);
            shifted = P3{xn, yn, m.layerType == LT_PIXEL ? up.p.z : lo.p.z};
            shiftedPhi = shifted.phi();
            shiftedRt = shifted.rt();
            if (m.layerType == LT_PIXEL) md.dphi = deltaPhi(lo.phi, shifted);
            else md.dphi = phiMpiPi(up.phi - shiftedPhi);
        } else {
            md.dphi = deltaPhi(lo.phi, up.p);
        }
        CUT(1, std::abs(md.dphi) < miniCut, std::abs(md.dphi), kNone, miniCut);
        if (tilted) {
            if (m.layerType == LT_PIXEL)
                md.dphichange = lo.rt < shiftedRt ? dPhiChange(lo.p, lo.phi, shifted) : dPhiChange(shifted, shiftedPhi, lo.p);
            else
                md.dphichange = shiftedRt < up.rt ? dPhiChange(shifted, shiftedPhi, up.p) : dPhiChange(up.p, up.phi, shifted);
        } else {
            md.dphichange = dPhiChange(lo.p, lo.phi, up.p);
        }
        CUT(2, std::abs(md.dphichange) < miniCut, std::abs(md.dphichange), kNone, miniCut);
        if (rec) {  // MiniDoublet.cc:259, 327-343, 346, 391-410
            rec[0] = md.dz; rec[1] = md.shiftedDz; rec[2] = md.dphi; rec[3] = md.dphichange;
            rec[4] = deltaPhi(lo.phi, up.p);
            rec[5] = (!tilted || lo.rt < up.rt) ? dPhiChange(lo.p, lo.phi, up.p) : dPhiChange(up.p, up.phi, lo.p);
            rec[6] = miniCut;
        }
        return true;
    }

    // MiniDoublet::runMiniDoubletDefaultAlgoEndcap (MiniDoublet.cc:449-684)
    bool mdEndcap(const HitD& lo, const HitD& up, const Slot& s, MDRec& md, Trace* tr) const {
        const ModInfo& m = s.info;
        md.dz = lo.p.z - up.p.z;
        float dz = md.dz;
        const float dzCut = 10.f;  // side() is never Endcap (A-4)
        CUT(10, std::abs(dz) < dzCut, std::abs(dz), kNone, dzCut);
        const float drtCut = m.type == TYPE_PS ? 2.f : 10.f;
        float drt = std::abs(lo.rt - up.rt);
        CUT(11, drt < drtCut, drt, kNone, drtCut);
        float xn, yn, zn;
        shiftStripHits(lo, up, s, xn, yn, zn);
        if (m.type == TYPE_PS && m.layerType != LT_PIXEL) {
            P3 loS{xn, yn, lo.p.z};
            md.dphi = phiMpiPi(up.phi - loS.phi());
        } else {
            P3 upS{xn, yn, up.p.z};
            md.dphi = deltaPhi(lo.phi, upS);
        }
        if (m.type == TYPE_PS) {
            md.shiftedDz = m.layerType == LT_PIXEL ? lo.p.z - zn : up.p.z - zn;
            dz = md.shiftedDz;
        }
        float miniCut = m.layerType == LT_PIXEL ? dPhiThreshold(lo, s, md.dphi, dz) : dPhiThreshold(up, s, md.dphi, dz);
        CUT(12, std::abs(md.dphi) < miniCut, std::abs(md.dphi), kNone, miniCut);
        float dzFrac = float(double(std::abs(dz)) / std::fabs(double(lo.p.z)));
        md.dphichange = md.dphi / dzFrac * (1.f + dzFrac);
        CUT(13, std::abs(md.dphichange) < miniCut, std::abs(md.dphichange), kNone, miniCut);
        if (rec) {  // MiniDoublet.cc:467, 563-597, 608, 644-645
            rec[0] = md.dz; rec[1] = md.shiftedDz; rec[2] = md.dphi; rec[3] = md.dphichange;
            rec[4] = deltaPhi(lo.phi, up.p);
            rec[5] = rec[4] / dzFrac * (1.f + dzFrac);
            rec[6] = miniCut;
        }
        return true;
    }

    // MiniDoublet ctor + runMiniDoubletDefaultAlgo (MiniDoublet.cc:28-81, 228-241)
    bool mdAlgo(int loIdx, int upIdx, MDRec& md, Trace* tr) const {
        const HitD& lo = hits[loIdx];
        const HitD& up = hits[upIdx];
        const Slot& s = slots[lo.slot];
        md.lo = loIdx; md.up = upIdx; md.slot = lo.slot;
        md.dz = md.shiftedDz = md.dphi = md.dphichange = 0;
        md.anchor = (s.info.type == TYPE_PS && s.info.layerType != LT_PIXEL) ? upIdx : loIdx;
        if (s.info.subdet == SUB_BARREL) return mdBarrel(lo, up, s, md, tr);
        return mdEndcap(lo, up, s, md, tr);
    }

    // ---------------------------------------------------------------- SG
    // Segment::dAlphaThreshold (Segment.cc:886-994)
    void dAlphaThreshold(const MDRec& inMD, const MDRec& outMD, float th[3]) const {
        static const float mulsB[6] = {0.0052f, 0.0038f, 0.0034f, 0.0034f, 0.0032f, 0.0034f};
        static const float mulsE[5] = {0.006f, 0.006f, 0.006f, 0.006f, 0.006f};
        const Slot& si = slots[inMD.slot];
        const Slot& so = slots[outMD.slot];
        const ModInfo& mi = si.info;
        const ModInfo& mo = so.info;
        float sdMuls = mi.subdet == SUB_BARREL ? mulsB[mi.layer - 1] * 3.f / kPtCut : mulsE[mi.layer - 1] * 3.f / kPtCut;
        const HitD& ai = hits[inMD.anchor];
        const HitD& ao = hits[outMD.anchor];
        float segmentDr = float(std::sqrt(sq(ao.p.y - ai.p.y) + sq(ao.p.x - ai.p.x)));
        const float dAlpha_Bfield = lm.asinF(fmin2(segmentDr * k2Rinv1GeVf / kPtCut, kSinAlphaMax));
        const bool inTilted = mi.subdet == SUB_BARREL && mi.side != SIDE_CENTER;
        const bool outTilted = mo.subdet == SUB_BARREL && mo.side != SIDE_CENTER;
        const float di = si.drdzStrip, dO = so.drdzStrip;
        const float inTilt = inTilted ? float(double(0.5f * kPixelPSZpitch * di) / std::sqrt(double(1.f + di * di)) / double(si.gap)) : 0.f;
        const float outTilt = outTilted ? float(double(0.5f * kPixelPSZpitch * dO) / std::sqrt(double(1.f + dO * dO)) / double(so.gap)) : 0.f;
        float miniDelta = si.gap;
        float lumIn, lumOut;
        if (mi.subdet == SUB_BARREL) lumIn = inTilt * dAlpha_Bfield;
        else lumIn = float(double(inMD.dphi) * 15.0 / double(mi.type == TYPE_PS ? inMD.shiftedDz : inMD.dz));
        if (mo.subdet == SUB_BARREL) lumOut = outTilt * dAlpha_Bfield;
        else lumOut = float(double(outMD.dphi) * 15.0 / double(mo.type == TYPE_PS ? outMD.shiftedDz : outMD.dz));
        float resIn = 0.02f / miniDelta * (mi.subdet == SUB_BARREL ? 1.0f : std::abs(ai.p.z / ai.rt));
        float resOut = 0.02f / miniDelta * (mo.subdet == SUB_BARREL ? 1.0f : std::abs(ao.p.z / ao.rt));
        float res = resIn + resOut;
        double base = sq(res) + sq(sdMuls);
        if (mi.subdet == SUB_BARREL && mi.side == SIDE_CENTER) th[0] = float(double(dAlpha_Bfield) + std::sqrt(base));
        else th[0] = float(double(dAlpha_Bfield) + std::sqrt(base + sq(lumIn)));
        if (mo.subdet == SUB_BARREL && mo.side == SIDE_CENTER) th[1] = float(double(dAlpha_Bfield) + std::sqrt(base));
        else th[1] = float(double(dAlpha_Bfield) + std::sqrt(base + sq(lumOut)));
        th[2] = float(double(dAlpha_Bfield) + std::sqrt(base));
    }

    bool sgAlphaCuts(const MDRec& inMD, const MDRec& outMD, float sgAlpha, Trace* tr) const {
        float th[3];
        dAlphaThreshold(inMD, outMD, th);
        if (rec)  // recovars dAlpha_* (Segment.cc:532-534, 801-803)
            for (int k = 0; k < 3; ++k) rec[19 + k] = th[k];
        float a = inMD.dphichange - sgAlpha;
        CUT(24, std::abs(a) < th[0], std::abs(a), kNone, th[0]);
        float b = outMD.dphichange - sgAlpha;
        CUT(25, std::abs(b) < th[1], std::abs(b), kNone, th[1]);
        float c = outMD.dphichange - inMD.dphichange;
        CUT(26, std::abs(c) < th[2], std::abs(c), kNone, th[2]);
        return true;
    }

    // Segment::runSegmentDefaultAlgoBarrel (Segment.cc:364-599)
    bool sgBarrel(const MDRec& inMD, const MDRec& outMD, SGRec& sg, Trace* tr) const {
        static const float mulsB[6] = {0.0052f, 0.0038f, 0.0034f, 0.0034f, 0.0032f, 0.0034f};
        const ModInfo& mi = slots[inMD.slot].info;
        const float sdMuls = mulsB[mi.layer - 1] * 3.f / kPtCut;
        const HitD& ai = hits[inMD.anchor];
        const HitD& ao = hits[outMD.anchor];
        float rtIn = ai.rt, rtOut = ao.rt, zIn = ai.p.z, zOut = ao.p.z;
        const float sdSlope = lm.asinF(fmin2(rtOut * k2Rinv1GeVf / kPtCut, kSinAlphaMax));
        const float sdPVoff = 0.1f / rtOut;
        const float dzDrtScale = lm.tanF(sdSlope) / sdSlope;
        const float zGeom = mi.layer <= 2 ? 2.f * kPixelPSZpitch : 2.f * kStrip2SZpitch;
        float zLo = zIn + (zIn - kDeltaZLum) * (rtOut / rtIn - 1.f) * (zIn > 0.f ? 1.f : dzDrtScale) - zGeom;
        float zHi = zIn + (zIn + kDeltaZLum) * (rtOut / rtIn - 1.f) * (zIn < 0.f ? 1.f : dzDrtScale) + zGeom;
        CUT(20, zOut >= zLo && zOut <= zHi, zOut, zLo, zHi);
        const float sdCut = float(double(sdSlope) + std::sqrt(double(sdMuls * sdMuls + sdPVoff * sdPVoff)));
        const float dPhi = deltaPhi(ai.phi, ao.p);
        CUT(21, std::abs(dPhi) <= sdCut, std::abs(dPhi), kNone, sdCut);
        sg.dphichange = dPhiChange(ai.p, ai.phi, ao.p);
        if (rec) {  // Segment.cc:443-460, 480-484, 501 (rt windows, dphi, min/max stay 0)
            const float v[19] = {rtIn, rtOut, zIn, zOut, zLo, zHi, 0.f, 0.f, 0.f, 0.f, 0.f, sg.dphichange, 0.f, 0.f,
                                 sdCut, sdSlope, sdMuls, sdPVoff, dPhi};
            for (int k = 0; k < 19; ++k) rec[k] = v[k];
        }
        CUT(22, std::abs(sg.dphichange) <= sdCut, std::abs(sg.dphichange), kNone, sdCut);
        return sgAlphaCuts(inMD, outMD, sg.dphichange, tr);
    }

    // Segment::runSegmentDefaultAlgoEndcap (Segment.cc:601-859)
    bool sgEndcap(const MDRec& inMD, const MDRec& outMD, SGRec& sg, Trace* tr) const {
        const Slot& si = slots[inMD.slot];
        const Slot& so = slots[outMD.slot];
        const ModInfo& mi = si.info;
        const ModInfo& mo = so.info;
        bool found = false;
        for (uint32_t d : g->connected(mi.detId))
            if (d == mo.detId) { found = true; break; }
        bool l2r123 = (mi.ring == 1 || mi.ring == 2 || mi.ring == 3) && mi.layer == 2;
        CUT(30, found || l2r123, kNone, kNone, kNone);
        const HitD& ai = hits[inMD.anchor];
        const HitD& ao = hits[outMD.anchor];
        bool outEC2S = mo.type == TYPE_2S && mo.subdet == SUB_ENDCAP;
        const HitD& aoHi = outEC2S ? edges[ao.hiEdge] : ao;
        const HitD& aoLo = outEC2S ? edges[ao.loEdge] : ao;
        float rtIn = ai.rt, rtOut = ao.rt, zIn = ai.p.z, zOut = ao.p.z;
        const float sdSlope = lm.asinF(fmin2(rtOut * k2Rinv1GeVf / kPtCut, kSinAlphaMax));
        const float rtGeom = (rtIn < 60.f && rtOut < 60.f) ? 2.f * kPixelPSZpitch
                             : ((rtIn < 60.f || rtOut < 60.f) ? (kPixelPSZpitch + kStrip2SZpitch) : 2.f * kStrip2SZpitch);
        CUT(31, !(zIn * zOut < 0), zIn * zOut, 0, kNone);
        const float dz = zOut - zIn;
        const float dLum = std::copysign(kDeltaZLum, zIn);
        const float drtDzScale = sdSlope / lm.tanF(sdSlope);
        float rtLo = fmax2(rtIn * (1.f + dz / (zIn + dLum) * drtDzScale) - rtGeom, rtIn - 0.5f * rtGeom);
        float rtHi = rtIn * (zOut - dLum) / (zIn - dLum) + rtGeom;
        CUT(32, rtOut >= rtLo && rtOut <= rtHi, rtOut, rtLo, rtHi);
        const float dPhiPos = deltaPhi(ai.phi, ao.p);
        const float sdCut = sdSlope;
        const float dHigh = deltaPhi(ai.phi, aoHi.p);
        const float dLow = deltaPhi(ai.phi, aoLo.p);
        bool hiBigger = absTrunc(dHigh) > absTrunc(dLow);  // A-3
        sg.dphi = dPhiPos;
        sg.dphimax = hiBigger ? dHigh : dLow;
        sg.dphimin = hiBigger ? dLow : dHigh;
        CUT(33, std::abs(dPhiPos) <= sdCut, std::abs(dPhiPos), kNone, sdCut);
        const float dzFrac = dz / zIn;
        sg.dphichange = dPhiPos / dzFrac * (1.f + dzFrac);
        sg.dphichangemin = sg.dphimin / dzFrac * (1.f + dzFrac);
        sg.dphichangemax = sg.dphimax / dzFrac * (1.f + dzFrac);
        if (rec) {  // Segment.cc:681-682, 721-722, 746-752, 769-771 (z windows stay 0; sdMuls, sdPVoff -999)
            const float v[19] = {rtIn, rtOut, 0.f, 0.f, 0.f, 0.f, rtLo, rtHi, sg.dphi, sg.dphimin, sg.dphimax,
                                 sg.dphichange, sg.dphichangemin, sg.dphichangemax, sdCut, sdSlope, -999.f, -999.f,
                                 dPhiPos};
            for (int k = 0; k < 19; ++k) rec[k] = v[k];
        }
        CUT(34, std::abs(sg.dphichange) <= sdCut, std::abs(sg.dphichange), kNone, sdCut);
        return sgAlphaCuts(inMD, outMD, sg.dphichange, tr);
    }

    // Segment::runSegmentDefaultAlgo dispatch (Segment.cc:337-362)
    bool sgAlgo(int mdIn, int mdOut, SGRec& sg, Trace* tr) const {
        sg.mdIn = mdIn; sg.mdOut = mdOut;
        sg.dphi = sg.dphimin = sg.dphimax = sg.dphichange = sg.dphichangemin = sg.dphichangemax = 0;
        const MDRec& a = mds[mdIn];
        const MDRec& b = mds[mdOut];
        bool inB = slots[a.slot].info.subdet == SUB_BARREL;
        bool outB = slots[b.slot].info.subdet == SUB_BARREL;
        if (inB) return outB ? sgBarrel(a, b, sg, tr) : sgEndcap(a, b, sg, tr);
        return outB ? sgBarrel(a, b, sg, tr) : sgEndcap(a, b, sg, tr);
    }

    // ---------------------------------------------------------------- TL
    // Tracklet::runDeltaBetaIterations (Tracklet.cc:3181-3324): the state the
    // cuts use and, with tk (a tracklet's map keys), the values it records.
    void deltaBetaIterations(float& betaIn, float& betaOut, float& betaAv, float& ptBeta, float sdInDr, float sdOutDr,
                             float dr, int lIn, float ptBetaMax, float* tk = nullptr) const {
        auto corr = [&](float len, float pt, float signOf) {
            return double(std::copysign(double(lm.asinF(fmin2(len * k2Rinv1GeVf / std::abs(pt), kSinAlphaMax))), double(signOf)));
        };
        auto ptOf = [&](float beta) { return float(double(dr * k2Rinv1GeVf) / std::sin(double(beta))); };
        // one iteration's record: betaIn_<n>, betaOut_<n>, betaAv_<n>, betaPt_<n>, dBeta_<n>
        auto put = [&](int in, int out, int av, int pt, int db, float bi, float bo, float ba, float bp) {
            if (!tk) return;
            tk[in] = bi; tk[out] = bo; tk[av] = ba; tk[pt] = bp; tk[db] = bi - bo;
        };
        auto put0 = [&](float bi, float bo, float ba, float bp) { put(TK_betaIn_0th, TK_betaOut_0th, TK_betaAv_0th, TK_betaPt_0th, TK_dBeta_0th, bi, bo, ba, bp); };
        auto put1 = [&](float bi, float bo, float ba, float bp) { put(TK_betaIn_1st, TK_betaOut_1st, TK_betaAv_1st, TK_betaPt_1st, TK_dBeta_1st, bi, bo, ba, bp); };
        auto put2 = [&](float bi, float bo, float ba, float bp) { put(TK_betaIn_2nd, TK_betaOut_2nd, TK_betaAv_2nd, TK_betaPt_2nd, TK_dBeta_2nd, bi, bo, ba, bp); };
        auto put3 = [&](float bi, float bo, float ba, float bp) { put(TK_betaIn_3rd, TK_betaOut_3rd, TK_betaAv_3rd, TK_betaPt_3rd, TK_dBeta_3rd, bi, bo, ba, bp); };
        auto put4 = [&](float bi, float bo, float ba, float bp) { put(TK_betaIn_4th, TK_betaOut_4th, TK_betaAv_4th, TK_betaPt_4th, TK_dBeta_4th, bi, bo, ba, bp); };
        if (betaIn * betaOut > 0.f &&
            (std::abs(ptBeta) < 4.f * ptBetaMax || (lIn >= 11 && std::abs(ptBeta) < 8.f * ptBetaMax))) {
            const float in0 = betaIn, out0 = betaOut;
            if (tk) {  // Tracklet.cc:3192-3200
                tk[TK_betacormode] = 1;
                put0(betaIn, betaOut, betaAv, ptBeta);
                tk[TK_betaIn_1stCorr] = float(corr(sdInDr, ptBeta, betaIn));
                tk[TK_betaOut_1stCorr] = float(corr(sdOutDr, ptBeta, betaOut));
            }
            const float inUpd = float(double(betaIn) + corr(sdInDr, ptBeta, betaIn));
            const float outUpd = float(double(betaOut) + corr(sdOutDr, ptBeta, betaOut));
            betaAv = 0.5f * (inUpd + outUpd);
            ptBeta = ptOf(betaAv);
            if (tk) {  // :3207-3213
                put1(inUpd, outUpd, betaAv, ptBeta);
                tk[TK_betaIn_2ndCorr] = float(corr(sdInDr, ptBeta, betaIn));
                tk[TK_betaOut_2ndCorr] = float(corr(sdOutDr, ptBeta, betaOut));
            }
            betaIn = float(double(betaIn) + corr(sdInDr, ptBeta, betaIn));
            betaOut = float(double(betaOut) + corr(sdOutDr, ptBeta, betaOut));
            betaAv = 0.5f * (betaIn + betaOut);
            ptBeta = ptOf(betaAv);
            if (tk) {  // :3221-3239: the 3rd restarts from the 0th values, the 4th from the 3rd
                put2(betaIn, betaOut, betaAv, ptBeta);
                tk[TK_betaIn_3rdCorr] = float(corr(sdInDr, ptBeta, betaIn));
                tk[TK_betaOut_3rdCorr] = float(corr(sdOutDr, ptBeta, betaOut));
                const float in3 = float(double(in0) + corr(sdInDr, ptBeta, betaIn));
                const float out3 = float(double(out0) + corr(sdOutDr, ptBeta, betaOut));
                const float av3 = 0.5f * (in3 + out3);
                const float pt3 = ptOf(av3);
                put3(in3, out3, av3, pt3);
                const float in4 = float(double(in0) + corr(sdInDr, pt3, in3));
                const float out4 = float(double(out0) + corr(sdOutDr, pt3, out3));
                const float av4 = 0.5f * (in4 + out4);
                put4(in4, out4, av4, ptOf(av4));
            }
        } else if (lIn < 11 && double(std::abs(betaOut)) < 0.2 * double(std::abs(betaIn)) &&
                   std::abs(ptBeta) < 12.f * ptBetaMax) {
            if (tk) { tk[TK_betacormode] = 2; put0(betaIn, betaOut, betaAv, ptBeta); }  // :3245-3250
            const float ptBetaIn = ptOf(betaIn);
            const float inUpd = float(double(betaIn) + corr(sdInDr, ptBetaIn, betaIn));
            const float outUpd = float(double(betaOut) + corr(sdOutDr, ptBetaIn, betaIn));
            betaAv = (std::abs(betaOut) > 0.2f * std::abs(betaIn)) ? (0.5f * (inUpd + outUpd)) : inUpd;
            ptBeta = ptOf(betaAv);
            betaIn = float(double(betaIn) + corr(sdInDr, ptBeta, betaIn));
            betaOut = float(double(betaOut) + corr(sdOutDr, ptBeta, betaIn));
            betaAv = 0.5f * (betaIn + betaOut);
            ptBeta = ptOf(betaAv);
            // :3263-3285: the 1st ... 4th all hold the final values
            put1(betaIn, betaOut, betaAv, ptBeta); put2(betaIn, betaOut, betaAv, ptBeta);
            put3(betaIn, betaOut, betaAv, ptBeta); put4(betaIn, betaOut, betaAv, ptBeta);
        } else if (tk) {  // :3290-3320: nothing changes
            tk[TK_betacormode] = 3;
            put0(betaIn, betaOut, betaAv, ptBeta); put1(betaIn, betaOut, betaAv, ptBeta);
            put2(betaIn, betaOut, betaAv, ptBeta); put3(betaIn, betaOut, betaAv, ptBeta);
            put4(betaIn, betaOut, betaAv, ptBeta);
        }
    }

    // MathUtil::getCenterFromThreePoints (MathUtil.cc:44-106): the centre of the
    // circle through A, B, C from the perpendicular bisectors of AB and BC; the
    // origin (a default Hit) when the two slopes are equal.  Returns the distance
    // from A to it, as TrackCandidate.cc:196-197 computes it.
    static void centerFromThree(const P3& A, const P3& B, const P3& C, float& cx, float& cy) {
        const float xA = A.x, yA = A.y, xB = B.x, yB = B.y, xC = C.x, yC = C.y;
        const float xMidAB = float(double(xA + xB) / 2.), yMidAB = float(double(yA + yB) / 2.);
        const bool infAB = (xB - xA) == 0, infBC = (xC - xB) == 0;
        const bool zeroAB = (yB - yA) == 0;
        const float sAB = infAB ? 0.f : (yB - yA) / (xB - xA);
        const float sBC = infBC ? 0.f : (yC - yB) / (xC - xB);
        const float perpAB = (infAB || zeroAB) ? 0.f : float(-1. / double(sAB));
        cx = 0.f;
        cy = 0.f;
        if (!((sAB - sBC) == 0)) {
            const float num = sAB * sBC * (yA - yC) + sBC * (xA + xB) - sAB * (xB + xC);
            cx = float(double(num) / (2. * double(sBC - sAB)));
            cy = perpAB * (cx - xMidAB) + yMidAB;
        }
    }
    static float radiusFromA(const P3& A, const P3& B, const P3& C) {
        float cx, cy;
        centerFromThree(A, B, C, cx, cy);
        return float(std::sqrt(sq(double(A.x - cx)) + sq(double(A.y - cy))));
    }

    // TrackCandidate::runTrackCandidateDefaultAlgo's recorded values (TrackCandidate.cc:183-207):
    // innerR through the inner TL's first three anchors, outerR through the outer
    // TL's last three, dR = (innerR - outerR) / innerR.
    void tcRecoVars(const PairRec& inner, const PairRec& outer, float* out) const {
        auto anchor = [&](int md) { return hits[mds[md].anchor].p; };
        const float innerR = radiusFromA(anchor(sgs[inner.in].mdIn), anchor(sgs[inner.in].mdOut), anchor(sgs[inner.out].mdIn));
        const float outerR = radiusFromA(anchor(sgs[outer.in].mdOut), anchor(sgs[outer.out].mdIn), anchor(sgs[outer.out].mdOut));
        out[0] = (innerR - outerR) / innerR;
        out[1] = innerR;
    }

    // The values the BBBB (_v2) and EEEE algorithms record alike at their end
    // (Tracklet.cc:1399-1493, 2522-2617): the four anchors' x, y and module
    // detIds (as float), the constants, and the midpoint dBeta.
    void tlCommonKeys(const SGRec& sIn, const SGRec& sOut, float* tk) const {
        const HitD* h[4] = {&hits[mds[sIn.mdIn].anchor], &hits[mds[sIn.mdOut].anchor], &hits[mds[sOut.mdIn].anchor],
                            &hits[mds[sOut.mdOut].anchor]};
        const int xk[4] = {TK_hit1_x, TK_hit2_x, TK_hit3_x, TK_hit4_x};
        const int yk[4] = {TK_hit1_y, TK_hit2_y, TK_hit3_y, TK_hit4_y};
        for (int k = 0; k < 4; ++k) { tk[xk[k]] = h[k]->p.x; tk[yk[k]] = h[k]->p.y; }
        tk[TK_innerSgInnerMdDetId] = float(slots[h[0]->slot].info.detId);
        tk[TK_innerSgOuterMdDetId] = float(slots[h[1]->slot].info.detId);
        tk[TK_outerSgInnerMdDetId] = float(slots[h[2]->slot].info.detId);
        tk[TK_outerSgOuterMdDetId] = float(slots[h[3]->slot].info.detId);
        tk[TK_sinAlphaMax] = kSinAlphaMax;
        tk[TK_deltaZLum] = kDeltaZLum;
        tk[TK_k2Rinv1GeVf] = k2Rinv1GeVf;
        tk[TK_kRinv1GeVf] = kRinv1GeVf;
        tk[TK_pixelPSZpitch] = kPixelPSZpitch;
        tk[TK_strip2SZpitch] = kStrip2SZpitch;
        tk[TK_ptCut] = kPtCut;
        tk[TK_betacormode] = 0;  // the algorithm's own betacormode, never changed from 0 (:1245, :2351)
        // dBeta_midPoint (Tracklet.cc:1409-1437): turning angles at the two segment midpoints
        const float x1 = h[0]->p.x, y1 = h[0]->p.y, x2 = h[1]->p.x, y2 = h[1]->p.y;
        const float x3 = h[2]->p.x, y3 = h[2]->p.y, x4 = h[3]->p.x, y4 = h[3]->p.y;
        const float inMidX = float(double(x1 + x2) / 2.), inMidY = float(double(y1 + y2) / 2.);
        const float outMidX = float(double(x3 + x4) / 2.), outMidY = float(double(y3 + y4) / 2.);
        const float ax = x2 - inMidX, ay = y2 - inMidY;
        const float bx = outMidX - inMidX, by = outMidY - inMidY;
        const float cx = x4 - outMidX, cy = y4 - outMidY;
        const float aMag = float(std::sqrt(double(ax * ax + ay * ay)));
        const float bMag = float(std::sqrt(double(bx * bx + by * by)));
        const float cMag = float(std::sqrt(double(cx * cx + cy * cy)));
        const float ab = ax * bx + ay * by, bc = bx * cx + by * cy;
        tk[TK_dBeta_midPoint] = lm.acosF(ab / aMag / bMag) - lm.acosF(bc / bMag / cMag);
    }

    // Shared tail of the three variants: rescale of the RH ranges and the dBeta cut
    // (Tracklet.cc:1344-1501, 1875-1990, 2377-2520).
    static void rescaleRH(float beta, float& mn, float& mx) {
        const float sf = (std::abs(mn + mx) > 0) ? float(2.f * beta / std::abs(mn + mx)) : float(0.);
        mn *= sf; mx *= sf;
    }

    // runTrackletDefaultAlgoBarrelBarrelBarrelBarrel_v2 (Tracklet.cc:902-1538)
    bool tlBBBB(const SGRec& sIn, const SGRec& sOut, Trace* tr) const {
        const MDRec& mInLo = mds[sIn.mdIn];
        const MDRec& mInUp = mds[sIn.mdOut];
        const MDRec& mOutLo = mds[sOut.mdIn];
        const MDRec& mOutUp = mds[sOut.mdOut];
        const HitD& hInLo = hits[mInLo.anchor];
        const HitD& hInUp = hits[mInUp.anchor];
        const HitD& hOutLo = hits[mOutLo.anchor];
        const HitD& hOutUp = hits[mOutUp.anchor];
        const bool isPS_InLo = slots[mInLo.slot].info.type == TYPE_PS;
        const bool isPS_OutLo = slots[mOutLo.slot].info.type == TYPE_PS;
        const float rt_InLo = hInLo.rt, rt_OutLo = hOutLo.rt, z_InLo = hInLo.p.z, z_OutLo = hOutLo.p.z;
        const float alpha1GeV_OutLo = lm.asinF(fmin2(rt_OutLo * k2Rinv1GeVf / kPtCut, kSinAlphaMax));
        const float rtRatio = rt_OutLo / rt_InLo;
        const float dzDrtScale = float(std::tan(double(alpha1GeV_OutLo)) / double(alpha1GeV_OutLo));
        const float zpitch_InLo = isPS_InLo ? kPixelPSZpitch : kStrip2SZpitch;
        const float zpitch_OutLo = isPS_OutLo ? kPixelPSZpitch : kStrip2SZpitch;
        const float zHi = z_InLo + (z_InLo + kDeltaZLum) * (rtRatio - 1.f) * (z_InLo < 0.f ? 1.f : dzDrtScale) + (zpitch_InLo + zpitch_OutLo);
        const float zLo = z_InLo + (z_InLo - kDeltaZLum) * (rtRatio - 1.f) * (z_InLo > 0.f ? 1.f : dzDrtScale) - (zpitch_InLo + zpitch_OutLo);
        float* tk = rec ? rec + kTLTyped : nullptr;
        if (tk) { tk[TK_z_OutLo] = z_OutLo; tk[TK_zLo] = zLo; tk[TK_zHi] = zHi; }  // Tracklet.cc:992-994
        CUT(40, z_OutLo >= zLo && z_OutLo <= zHi, z_OutLo, zLo, zHi);
        const float drt_OutLo_InLo = rt_OutLo - rt_InLo;
        const float r3_InLo = float(std::sqrt(double(z_InLo * z_InLo + rt_InLo * rt_InLo)));
        const float drt_InSeg = hInUp.rt - hInLo.rt;
        const float dz_InSeg = hInUp.p.z - hInLo.p.z;
        const float dr3_InSeg = hInUp.r3 - hInLo.r3;
        const float coshEta = dr3_InSeg / drt_InSeg;
        float dzErr = (zpitch_InLo + zpitch_OutLo) * (zpitch_InLo + zpitch_OutLo) * 2.f;
        const float sdlThetaMulsF = float(double(0.015f) * std::sqrt(double(0.1f) + 0.2 * double(rt_OutLo - rt_InLo) / double(50.f)) *
                                          std::sqrt(double(r3_InLo / rt_InLo)));
        const float sdlMuls = sdlThetaMulsF * 3.f / kPtCut * 4.f;
        dzErr += sdlMuls * sdlMuls * drt_OutLo_InLo * drt_OutLo_InLo / 3.f * coshEta * coshEta;
        dzErr = float(std::sqrt(double(dzErr)));
        const float dzMean = dz_InSeg / drt_InSeg * drt_OutLo_InLo;
        const float zWindow = dzErr / drt_InSeg * drt_OutLo_InLo + (zpitch_InLo + zpitch_OutLo);
        const float zLoP = z_InLo + dzMean * (z_InLo > 0.f ? 1.f : dzDrtScale) - zWindow;
        const float zHiP = z_InLo + dzMean * (z_InLo < 0.f ? 1.f : dzDrtScale) + zWindow;
        if (tk) { tk[TK_zLoPointed] = zLoP; tk[TK_zHiPointed] = zHiP; }  // :1066-1067
        CUT(41, z_OutLo >= zLoP && z_OutLo <= zHiP, z_OutLo, zLoP, zHiP);
        const float sdlPVoff = 0.1f / rt_OutLo;
        const float sdlCut = float(double(alpha1GeV_OutLo) + std::sqrt(double(sdlMuls * sdlMuls + sdlPVoff * sdlPVoff)));
        const float deltaPhiPos = deltaPhi(hInUp.phi, hOutUp.p);
        if (tk) { tk[TK_deltaPhiPos] = deltaPhiPos; tk[TK_sdlCut] = sdlCut; }  // :1098-1099
        CUT(42, std::abs(deltaPhiPos) <= sdlCut, std::abs(deltaPhiPos), kNone, sdlCut);
        const P3 mid = midpoint(hInLo.p, hOutLo.p);
        const float dPhi = deltaPhi(mid.phi(), sub(hOutLo.p, hInLo.p));
        if (tk) tk[TK_dPhi] = dPhi;  // :1152
        CUT(43, std::abs(dPhi) <= sdlCut, std::abs(dPhi), kNone, sdlCut);
        const float alpha_InLo = sIn.dphichange;
        const ModInfo& mOU = slots[mOutUp.slot].info;
        const bool isEC_lastLayer = mOU.subdet == SUB_ENDCAP && mOU.type == TYPE_2S;
        // GeometryUtil::stripHigh/LowEdgeHit recomputed on the anchor (Tracklet.cc:1171-1172)
        const HitD hiE = isEC_lastLayer ? edgeHit(hOutUp, true) : hOutUp;
        const HitD loE = isEC_lastLayer ? edgeHit(hOutUp, false) : hOutUp;
        const float alpha_OutUp = deltaPhi(hOutUp.phi, sub(hOutUp.p, hOutLo.p));
        const float alpha_OutUp_hi = deltaPhi(hiE.phi, sub(hiE.p, hOutLo.p));
        const float alpha_OutUp_lo = deltaPhi(loE.phi, sub(loE.p, hOutLo.p));
        const P3 axis = sub(hOutUp.p, hInLo.p);
        const P3 axisHi = sub(hiE.p, hInLo.p);
        const P3 axisLo = sub(loE.p, hInLo.p);
        float betaIn = alpha_InLo - deltaPhi(hInLo.phi, axis);
        float betaInRHmin = betaIn, betaInRHmax = betaIn;
        float betaOut = -alpha_OutUp + deltaPhi(hOutUp.phi, axis);
        float betaOutRHmin = -alpha_OutUp_hi + deltaPhi(hiE.phi, axisHi);
        float betaOutRHmax = -alpha_OutUp_lo + deltaPhi(loE.phi, axisLo);
        const float drt_tl_axis = axis.rt();
        const float rt_InSeg = sub(hInUp.p, hInLo.p).rt();
        const float betaIn_cut = lm.asinF(fmin2((-rt_InSeg * 1.f + drt_tl_axis) * k2Rinv1GeVf / kPtCut, kSinAlphaMax)) + (0.02f / drt_InSeg);
        CUT(44, std::abs(betaInRHmin) < betaIn_cut, std::abs(betaInRHmin), kNone, betaIn_cut);
        if (rec) { rec[0] = betaInRHmin; rec[1] = betaIn_cut; }  // setBetaIn, setBetaInCut (Tracklet.cc:1218-1219)
        if (tk) {  // :1197-1204, 1217 (the RH ranges are recorded again, rescaled, at the end)
            tk[TK_betaIn_0th] = betaIn; tk[TK_betaOut_0th] = betaOut;
            tk[TK_rawBetaIn] = betaIn; tk[TK_rawBetaOut] = betaOut;
            tk[TK_betaIn_cut] = betaIn_cut;
        }
        float betaAv = 0.5f * (betaIn + betaOut);
        float pt_beta = float(double(drt_tl_axis * k2Rinv1GeVf) / std::sin(double(betaAv)));
        const float pt_betaMax = 7.0f;
        const int lIn = 5, lOut = isEC_lastLayer ? 11 : 5;
        const float sdOut_dr = sub(hOutUp.p, hOutLo.p).rt();
        const float sdOut_d = hOutUp.rt - hOutLo.rt;
        deltaBetaIterations(betaIn, betaOut, betaAv, pt_beta, rt_InSeg, sdOut_dr, drt_tl_axis, lIn, pt_betaMax, tk);
        rescaleRH(betaIn, betaInRHmin, betaInRHmax);
        rescaleRH(betaOut, betaOutRHmin, betaOutRHmax);
        const float dBetaMuls = sdlThetaMulsF * 4.f / fmin2(std::abs(pt_beta), pt_betaMax);
        const float alpha_OutLo = sOut.dphichange;
        const float alphaInAbsReg = fmax2(std::abs(alpha_InLo), lm.asinF(fmin2(hInLo.rt * k2Rinv1GeVf / 3.0f, kSinAlphaMax)));
        const float alphaOutAbsReg = fmax2(std::abs(alpha_OutLo), lm.asinF(fmin2(hOutLo.rt * k2Rinv1GeVf / 3.0f, kSinAlphaMax)));
        const float dBetaInLum = lIn < 11 ? 0.0f : std::abs(alphaInAbsReg * kDeltaZLum / hInLo.p.z);
        const float dBetaOutLum = lOut < 11 ? 0.0f : std::abs(alphaOutAbsReg * kDeltaZLum / hOutLo.p.z);
        const float dBetaLum2 = (dBetaInLum + dBetaOutLum) * (dBetaInLum + dBetaOutLum);
        const float sinDPhi = lm.sinF(dPhi);
        const float dBetaROut2 = float(sq(double((hiE.rt - loE.rt) * sinDPhi / drt_tl_axis)));
        const float betaOut_cut = float(double(lm.asinF(fmin2(drt_tl_axis * k2Rinv1GeVf / kPtCut, kSinAlphaMax)) + (0.02f / sdOut_d)) +
                                        std::sqrt(double(dBetaLum2 + dBetaMuls * dBetaMuls)));
        CUT(45, std::abs(betaOut) < betaOut_cut, std::abs(betaOut), kNone, betaOut_cut);
        if (rec) { rec[2] = betaOut; rec[3] = betaOut_cut; }  // Tracklet.cc:1375-1376
        const float dBetaRes = 0.02f / fmin2(sdOut_d, drt_InSeg);
        const float dBetaCut2 = float(double(dBetaRes * dBetaRes * 2.0f + dBetaMuls * dBetaMuls + dBetaLum2 + 0.f + dBetaROut2) +
                                      0.25 * sq(double(std::abs(betaInRHmin - betaInRHmax) + std::abs(betaOutRHmin - betaOutRHmax))));
        const float dBeta = betaIn - betaOut;
        if (!tlNm1) CUT(46, dBeta * dBeta <= dBetaCut2, dBeta * dBeta, kNone, dBetaCut2);
        if (rec) { rec[4] = dBeta; rec[5] = std::sqrt(dBetaCut2); }  // Tracklet.cc:1498-1499
        if (tk) {  // :1438-1494
            tlCommonKeys(sIn, sOut, tk);
            tk[TK_betaIn] = betaIn; tk[TK_betaInRHmin] = betaInRHmin; tk[TK_betaInRHmax] = betaInRHmax;
            tk[TK_betaOut] = betaOut; tk[TK_betaOutRHmin] = betaOutRHmin; tk[TK_betaOutRHmax] = betaOutRHmax;
            tk[TK_dBeta] = dBeta; tk[TK_dBetaCut2] = dBetaCut2; tk[TK_dBetaLum2] = dBetaLum2;
            tk[TK_dBetaMuls] = dBetaMuls; tk[TK_dBetaRIn2] = 0.f; tk[TK_dBetaROut2] = dBetaROut2;
            tk[TK_dBetaRes] = dBetaRes; tk[TK_drt_tl_axis] = drt_tl_axis; tk[TK_dzDrtScale] = dzDrtScale;
            tk[TK_pt_betaIn] = float(double(drt_tl_axis * k2Rinv1GeVf) / std::sin(double(betaIn)));  // :1393-1394
            tk[TK_pt_betaOut] = float(double(drt_tl_axis * k2Rinv1GeVf) / std::sin(double(betaOut)));
            tk[TK_rt_InLo] = rt_InLo; tk[TK_rt_OutLo] = rt_OutLo; tk[TK_rtRatio_OutLoInLo] = rtRatio;
            tk[TK_drt_InSeg] = drt_InSeg; tk[TK_sdOut_d] = sdOut_d; tk[TK_alpha1GeV_OutLo] = alpha1GeV_OutLo;
            tk[TK_zpitch_InLo] = zpitch_InLo; tk[TK_zpitch_OutLo] = zpitch_OutLo; tk[TK_z_InLo] = z_InLo;
            tk[TK_pt_beta] = pt_beta; tk[TK_betaAv] = betaAv;
        }
        return true;
    }

    // runTrackletDefaultAlgoBarrelBarrelEndcapEndcap (Tracklet.cc:1540-2016)
    bool tlBBEE(const SGRec& sIn, const SGRec& sOut, Trace* tr) const {
        const MDRec& mInLo = mds[sIn.mdIn];
        const MDRec& mInUp = mds[sIn.mdOut];
        const MDRec& mOutLo = mds[sOut.mdIn];
        const MDRec& mOutUp = mds[sOut.mdOut];
        const HitD& hInLo = hits[mInLo.anchor];
        const HitD& hInUp = hits[mInUp.anchor];
        const HitD& hOutLo = hits[mOutLo.anchor];
        const HitD& hOutUp = hits[mOutUp.anchor];
        const bool psInLo = slots[mInLo.slot].info.type == TYPE_PS;
        const bool psOutLo = slots[mOutLo.slot].info.type == TYPE_PS;
        const float zGeom = (psInLo ? kPixelPSZpitch : kStrip2SZpitch) + (psOutLo ? kPixelPSZpitch : kStrip2SZpitch);
        const float rtIn = hInLo.rt, rtOut = hOutLo.rt, zIn = hInLo.p.z, zOut = hOutLo.p.z;
        const float sdlSlope = lm.asinF(fmin2(rtOut * k2Rinv1GeVf / kPtCut, kSinAlphaMax));
        const float dzDrtScale = float(std::tan(double(sdlSlope)) / double(sdlSlope));
        CUT(50, zIn * zOut > 0, zIn * zOut, 0, kNone);
        const float dLum = std::copysign(kDeltaZLum, zIn);
        const float rtGeom1 = psOutLo ? kPixelPSZpitch : kStrip2SZpitch;
        const float zGeom1 = std::copysign(zGeom, zIn);
        const float rtLo = rtIn * (1.f + (zOut - zIn - zGeom1) / (zIn + zGeom1 + dLum) / dzDrtScale) - rtGeom1;
        CUT(51, rtOut >= rtLo, rtOut, rtLo, kNone);
        float zInForHi = zIn - zGeom1 - dLum;
        if (zInForHi * zIn < 0) zInForHi = std::copysign(0.1f, zIn);
        const float rtHi = rtIn * (1.f + (zOut - zIn + zGeom1) / zInForHi) + rtGeom1;
        CUT(52, rtOut >= rtLo && rtOut <= rtHi, rtOut, rtLo, rtHi);
        const float rIn = float(std::sqrt(double(zIn * zIn + rtIn * rtIn)));
        const float drtSDIn = hInUp.rt - hInLo.rt;
        const float dzSDIn = hInUp.p.z - hInLo.p.z;
        const float dr3SDIn = hInUp.r3 - hInLo.r3;
        const float coshEta = dr3SDIn / drtSDIn;
        const float dzOutInAbs = std::abs(zOut - zIn);
        const float multDzDr = dzOutInAbs * coshEta / (coshEta * coshEta - 1.f);
        const float kZ = (zOut - zIn) / dzSDIn;
        const float sdlThetaMulsF = float(double(0.015f) * std::sqrt(double(0.1f) + 0.2 * double(rtOut - rtIn) / double(50.f)) *
                                          std::sqrt(double(rIn / rtIn)));
        const float sdlMuls = sdlThetaMulsF * 3.f / kPtCut * 4.f;
        CUT(53, kZ >= 0 && rtOut >= rtLo && rtOut <= rtHi, kZ, 0, kNone);
        const float sdlPVoff = 0.1f / rtOut;
        const float sdlCut = float(double(sdlSlope) + std::sqrt(double(sdlMuls * sdlMuls + sdlPVoff * sdlPVoff)));
        const float deltaPhiPos = deltaPhi(hInUp.phi, hOutUp.p);
        CUT(54, std::abs(deltaPhiPos) <= sdlCut, std::abs(deltaPhiPos), kNone, sdlCut);
        const P3 mid = midpoint(hInLo.p, hOutLo.p);
        const float dPhi = deltaPhi(mid.phi(), sub(hOutLo.p, hInLo.p));
        CUT(55, std::abs(dPhi) <= sdlCut, std::abs(dPhi), kNone, sdlCut);
        const float sdIn_alpha = sIn.dphichange;
        const float sdOut_alpha = sIn.dphichange;  // A-8
        const float sdOut_alphaOut = deltaPhi(hOutUp.phi, sub(hOutUp.p, hOutLo.p));
        const float sdOut_alphaOut_min = phiMpiPi(sOut.dphichangemin - sOut.dphimin);
        const float sdOut_alphaOut_max = phiMpiPi(sOut.dphichangemax - sOut.dphimax);
        const P3 dr3 = sub(hOutUp.p, hInLo.p);
        float betaIn = sdIn_alpha - deltaPhi(hInLo.phi, dr3);
        float betaOut = -sdOut_alphaOut + deltaPhi(hOutUp.phi, dr3);
        float betaInRHmin = betaIn, betaInRHmax = betaIn;  // the 2S test never holds (A-9)
        float betaOutRHmin = betaOut - sdOut_alphaOut_min + sdOut_alphaOut;
        float betaOutRHmax = betaOut - sdOut_alphaOut_max + sdOut_alphaOut;
        if (std::abs(betaOutRHmin) > std::abs(betaOutRHmax)) std::swap(betaOutRHmax, betaOutRHmin);
        if (std::abs(betaInRHmin) > std::abs(betaInRHmax)) std::swap(betaInRHmax, betaInRHmin);
        const float sdIn_dr = sub(hInUp.p, hInLo.p).rt();
        const float sdIn_d = hInUp.rt - hInLo.rt;
        const float dr = dr3.rt();
        const float betaIn_cut = lm.asinF(fmin2((-sdIn_dr * 1.f + dr) * k2Rinv1GeVf / kPtCut, kSinAlphaMax)) + (0.02f / sdIn_d);
        CUT(56, std::abs(betaInRHmin) < betaIn_cut, std::abs(betaInRHmin), kNone, betaIn_cut);
        float betaAv = 0.5f * (betaIn + betaOut);
        float pt_beta = float(double(dr * k2Rinv1GeVf) / std::sin(double(betaAv)));
        const float pt_betaMax = 7.0f;
        const int lIn = 5, lOut = 11;
        const float sdOut_dr = sub(hOutUp.p, hOutLo.p).rt();
        const float sdOut_d = hOutUp.rt - hOutLo.rt;
        // the only map keys this variant writes (Tracklet.cc:1885)
        deltaBetaIterations(betaIn, betaOut, betaAv, pt_beta, sdIn_dr, sdOut_dr, dr, lIn, pt_betaMax,
                            rec ? rec + kTLTyped : nullptr);
        rescaleRH(betaIn, betaInRHmin, betaInRHmax);
        rescaleRH(betaOut, betaOutRHmin, betaOutRHmax);
        const float dBetaMuls = sdlThetaMulsF * 4.f / fmin2(std::abs(pt_beta), pt_betaMax);
        const float alphaInAbsReg = fmax2(std::abs(sdIn_alpha), lm.asinF(fmin2(hInLo.rt * k2Rinv1GeVf / 3.0f, kSinAlphaMax)));
        const float alphaOutAbsReg = fmax2(std::abs(sdOut_alpha), lm.asinF(fmin2(hOutLo.rt * k2Rinv1GeVf / 3.0f, kSinAlphaMax)));
        const float dBetaInLum = lIn < 11 ? 0.0f : std::abs(alphaInAbsReg * kDeltaZLum / hInLo.p.z);
        const float dBetaOutLum = lOut < 11 ? 0.0f : std::abs(alphaOutAbsReg * kDeltaZLum / hOutLo.p.z);
        const float dBetaLum2 = (dBetaInLum + dBetaOutLum) * (dBetaInLum + dBetaOutLum);
        const float sinDPhi = lm.sinF(dPhi);
        const bool out2S = slots[mOutUp.slot].info.type == TYPE_2S;
        const HitD& hiE = out2S ? edges[hOutUp.hiEdge] : hOutUp;
        const HitD& loE = out2S ? edges[hOutUp.loEdge] : hOutUp;
        const float dBetaROut2 = float(sq(double((hiE.rt - loE.rt) * sinDPhi / dr)));
        const float betaOut_cut = float(double(lm.asinF(fmin2(dr * k2Rinv1GeVf / kPtCut, kSinAlphaMax)) + (0.02f / sdOut_d)) +
                                        std::sqrt(double(dBetaLum2 + dBetaMuls * dBetaMuls)));
        CUT(57, std::abs(betaOut) < betaOut_cut, std::abs(betaOut), kNone, betaOut_cut);
        const float dBetaRes = 0.02f / fmin2(sdOut_d, sdIn_d);
        const float dBetaCut2 = float(double(dBetaRes * dBetaRes * 2.0f + dBetaMuls * dBetaMuls + dBetaLum2 + 0.f + dBetaROut2) +
                                      0.25 * sq(double(std::abs(betaInRHmin - betaInRHmax) + std::abs(betaOutRHmin - betaOutRHmax))));
        const float dBeta = betaIn - betaOut;
        if (!tlNm1) CUT(58, dBeta * dBeta <= dBetaCut2, dBeta * dBeta, kNone, dBetaCut2);
        if (rec) rec[4] = dBeta;  // the only typed value this variant sets (Tracklet.cc:1977); the rest stay 0
        return true;
    }

    // runTrackletDefaultAlgoEndcapEndcapEndcapEndcap (Tracklet.cc:2018-2665)
    bool tlEEEE(const SGRec& sIn, const SGRec& sOut, Trace* tr) const {
        const MDRec& mInLo = mds[sIn.mdIn];
        const MDRec& mInUp = mds[sIn.mdOut];
        const MDRec& mOutLo = mds[sOut.mdIn];
        const MDRec& mOutUp = mds[sOut.mdOut];
        const HitD& hInLo = hits[mInLo.anchor];
        const HitD& hInUp = hits[mInUp.anchor];
        const HitD& hOutLo = hits[mOutLo.anchor];
        const HitD& hOutUp = hits[mOutUp.anchor];
        const bool psInLo = slots[mInLo.slot].info.type == TYPE_PS;
        const bool psOutLo = slots[mOutLo.slot].info.type == TYPE_PS;
        const bool psInUp = slots[mInUp.slot].info.type == TYPE_PS;
        const float rtIn = hInLo.rt, rtOut = hOutLo.rt, zIn = hInLo.p.z, zOut = hOutLo.p.z;
        const float sdlSlope = lm.asinF(fmin2(rtOut * k2Rinv1GeVf / kPtCut, kSinAlphaMax));
        const float dzDrtScale = float(std::tan(double(sdlSlope)) / double(sdlSlope));
        CUT(60, zIn * zOut > 0, zIn * zOut, 0, kNone);
        const float dLum = std::copysign(kDeltaZLum, zIn);
        const float rtGeom = (psInLo && psOutLo) ? 2.f * kPixelPSZpitch
                             : (psInLo || psOutLo) ? (kPixelPSZpitch + kStrip2SZpitch) : 2.f * kStrip2SZpitch;
        const float dz = zOut - zIn;
        const float rtLo = rtIn * (1.f + dz / (zIn + dLum) / dzDrtScale) - rtGeom;
        CUT(61, rtOut >= rtLo, rtOut, rtLo, kNone);
        const float rtHi = rtIn * (1.f + dz / (zIn - dLum)) + rtGeom;
        CUT(62, rtOut >= rtLo && rtOut <= rtHi, rtOut, rtLo, rtHi);
        const float drtSDIn = hInUp.rt - hInLo.rt;
        const float dzSDIn = hInUp.p.z - hInLo.p.z;
        const float dr3SDIn = hInUp.r3 - hInLo.r3;
        const float coshEta = dr3SDIn / drtSDIn;
        const float dzOutInAbs = std::abs(zOut - zIn);
        const float multDzDr = dzOutInAbs * coshEta / (coshEta * coshEta - 1.f);
        const float kZ = (zOut - zIn) / dzSDIn;
        const float sdlThetaMulsF = float(double(0.015f) * std::sqrt(double(0.1f) + 0.2 * double(rtOut - rtIn) / double(50.f)));
        const float sdlMuls = sdlThetaMulsF * 3.f / kPtCut * 4.f;
        float drtErr = kPixelPSZpitch * kPixelPSZpitch * 2.f / dzSDIn / dzSDIn * dzOutInAbs * dzOutInAbs;
        drtErr += sdlMuls * sdlMuls * multDzDr * multDzDr / 3.f * coshEta * coshEta;
        drtErr = float(std::sqrt(double(drtErr)));
        const float drtMean = drtSDIn * dzOutInAbs / std::abs(dzSDIn);
        const float rtWindow = drtErr + rtGeom;
        const float rtLoP = rtIn + drtMean / dzDrtScale - rtWindow;
        const float rtHiP = rtIn + drtMean + rtWindow;
        if (psInLo && psInUp) CUT(63, kZ >= 0 && rtOut >= rtLoP && rtOut <= rtHiP, rtOut, rtLoP, rtHiP);
        const float sdlPVoff = 0.1f / rtOut;
        const float sdlCut = float(double(sdlSlope) + std::sqrt(double(sdlMuls * sdlMuls + sdlPVoff * sdlPVoff)));
        const float deltaPhiPos = deltaPhi(hInUp.phi, hOutUp.p);
        CUT(64, std::abs(deltaPhiPos) <= sdlCut, std::abs(deltaPhiPos), kNone, sdlCut);
        const P3 mid = midpoint(hInLo.p, hOutLo.p);
        const float dPhi = deltaPhi(mid.phi(), sub(hOutLo.p, hInLo.p));
        CUT(65, std::abs(dPhi) <= sdlCut, std::abs(dPhi), kNone, sdlCut);
        const float sdIn_alpha = sIn.dphichange;
        const float sdOut_alpha = sIn.dphichange;  // A-8
        const float sdOut_dPhiPos = deltaPhi(hOutLo.phi, hOutUp.p);
        const float sdOut_alphaOutRHmin = phiMpiPi(sOut.dphichangemin - sdOut_dPhiPos);
        const float sdOut_alphaOutRHmax = phiMpiPi(sOut.dphichangemax - sdOut_dPhiPos);
        const float sdOut_alphaOut = phiMpiPi(sOut.dphichange - sdOut_dPhiPos);
        const P3 dr3 = sub(hOutUp.p, hInLo.p);
        float betaIn = sdIn_alpha - deltaPhi(hInLo.phi, dr3);
        float betaOut = -sdOut_alphaOut + deltaPhi(hOutUp.phi, dr3);
        float betaInRHmin = betaIn + sIn.dphichangemin - sdIn_alpha;
        float betaInRHmax = betaIn + sIn.dphichangemax - sdIn_alpha;
        if (std::abs(betaInRHmin) > std::abs(betaInRHmax)) std::swap(betaInRHmax, betaInRHmin);
        float betaOutRHmin = betaOut - sdOut_alphaOutRHmin + sdOut_alphaOut;
        float betaOutRHmax = betaOut - sdOut_alphaOutRHmax + sdOut_alphaOut;
        if (std::abs(betaOutRHmin) > std::abs(betaOutRHmax)) std::swap(betaOutRHmax, betaOutRHmin);
        const float sdIn_dr = sub(hInUp.p, hInLo.p).rt();
        const float sdIn_d = hInUp.rt - hInLo.rt;
        const float dr = dr3.rt();
        // the betaIn cut is computed but never rejects (A-11); its value has no asin (A-12)
        if (rec) { rec[0] = betaInRHmin; rec[1] = (-sdIn_dr * 1.f + dr) * k2Rinv1GeVf / kPtCut + (0.02f / sdIn_d); }  // 2319-2324
        float* tk = rec ? rec + kTLTyped : nullptr;
        if (tk) {  // Tracklet.cc:2303-2308
            tk[TK_sdIn_alpha] = sdIn_alpha; tk[TK_sdOut_alphaOut] = sdOut_alphaOut;
            tk[TK_rawBetaInCorrection] = deltaPhi(hInLo.phi, dr3);
            tk[TK_rawBetaOutCorrection] = deltaPhi(hOutUp.phi, dr3);
            tk[TK_rawBetaIn] = betaIn; tk[TK_rawBetaOut] = betaOut;
        }
        float betaAv = 0.5f * (betaIn + betaOut);
        float pt_beta = float(double(dr * k2Rinv1GeVf) / std::sin(double(betaAv)));
        const float pt_betaMax = 7.0f;
        const int lIn = 11, lOut = 13;
        const float sdOut_dr = sub(hOutUp.p, hOutLo.p).rt();
        const float sdOut_d = hOutUp.rt - hOutLo.rt;
        deltaBetaIterations(betaIn, betaOut, betaAv, pt_beta, sdIn_dr, sdOut_dr, dr, lIn, pt_betaMax, tk);
        rescaleRH(betaIn, betaInRHmin, betaInRHmax);
        rescaleRH(betaOut, betaOutRHmin, betaOutRHmax);
        const float dBetaMuls = sdlThetaMulsF * 4.f / fmin2(std::abs(pt_beta), pt_betaMax);
        const float alphaInAbsReg = fmax2(std::abs(sdIn_alpha), lm.asinF(fmin2(hInLo.rt * k2Rinv1GeVf / 3.0f, kSinAlphaMax)));
        const float alphaOutAbsReg = fmax2(std::abs(sdOut_alpha), lm.asinF(fmin2(hOutLo.rt * k2Rinv1GeVf / 3.0f, kSinAlphaMax)));
        const float dBetaInLum = lIn < 11 ? 0.0f : std::abs(alphaInAbsReg * kDeltaZLum / hInLo.p.z);
        const float dBetaOutLum = lOut < 11 ? 0.0f : std::abs(alphaOutAbsReg * kDeltaZLum / hOutLo.p.z);
        const float dBetaLum2 = (dBetaInLum + dBetaOutLum) * (dBetaInLum + dBetaOutLum);
        const float betaOut_cut = float(double(lm.asinF(fmin2(dr * k2Rinv1GeVf / kPtCut, kSinAlphaMax)) + (0.02f / sdOut_d)) +
                                        std::sqrt(double(dBetaLum2 + dBetaMuls * dBetaMuls)));
        CUT(67, std::abs(betaOut) < betaOut_cut, std::abs(betaOut), kNone, betaOut_cut);
        if (rec) { rec[2] = betaOut; rec[3] = betaOut_cut; }  // Tracklet.cc:2499-2500
        const float dBetaRes = 0.02f / fmin2(sdOut_d, sdIn_d);
        const float dBetaCut2 = float(double(dBetaRes * dBetaRes * 2.0f + dBetaMuls * dBetaMuls + dBetaLum2 + 0.f + 0.f) +
                                      0.25 * sq(double(std::abs(betaInRHmin - betaInRHmax) + std::abs(betaOutRHmin - betaOutRHmax))));
        const float dBeta = betaIn - betaOut;
        if (!tlNm1) CUT(68, dBeta * dBeta <= dBetaCut2, dBeta * dBeta, kNone, dBetaCut2);
        if (rec) { rec[4] = dBeta; rec[5] = std::sqrt(dBetaCut2); }  // Tracklet.cc:2621-2622
        if (tk) {  // :2576-2617
            tlCommonKeys(sIn, sOut, tk);
            tk[TK_betaIn] = betaIn; tk[TK_betaInRHmin] = betaInRHmin; tk[TK_betaInRHmax] = betaInRHmax;
            tk[TK_betaOut] = betaOut; tk[TK_betaOutRHmin] = betaOutRHmin; tk[TK_betaOutRHmax] = betaOutRHmax;
            tk[TK_dBeta] = dBeta; tk[TK_dBetaCut2] = dBetaCut2; tk[TK_dBetaLum2] = dBetaLum2;
            tk[TK_dBetaMuls] = dBetaMuls; tk[TK_dBetaRIn2] = 0.f; tk[TK_dBetaROut2] = 0.f;
            tk[TK_dBetaRes] = dBetaRes; tk[TK_dr] = dr; tk[TK_dzDrtScale] = dzDrtScale;
            tk[TK_pt_betaIn] = float(double(dr * k2Rinv1GeVf) / std::sin(double(betaIn)));  // :2516-2517
            tk[TK_pt_betaOut] = float(double(dr * k2Rinv1GeVf) / std::sin(double(betaOut)));
            const float rtOut_o_rtIn = rtOut / rtIn;
            const float zGeom = (psInLo ? kPixelPSZpitch : kStrip2SZpitch) + (psOutLo ? kPixelPSZpitch : kStrip2SZpitch);
            tk[TK_rtIn] = rtIn; tk[TK_rtOut] = rtOut; tk[TK_rtOut_o_rtIn] = rtOut_o_rtIn;
            tk[TK_sdIn_d] = sdIn_d; tk[TK_sdOut_d] = sdOut_d; tk[TK_sdlSlope] = sdlSlope; tk[TK_zGeom] = zGeom;
            tk[TK_zIn] = zIn; tk[TK_zOut] = zOut;
            // :2127, the z window of the barrel algorithm, computed and recorded but not cut on
            tk[TK_zLo] = zIn + (zIn - kDeltaZLum) * (rtOut_o_rtIn - 1.f) * (zIn > 0.f ? 1.f : dzDrtScale) - zGeom;
            tk[TK_pt_beta] = pt_beta; tk[TK_betaAv] = betaAv;
        }
        return true;
    }

    // Tracklet::runTrackletDefaultAlgo dispatch (Tracklet.cc:144-296)
    bool tlAlgo(const SGRec& sIn, const SGRec& sOut, Trace* tr) const {
        bool b0 = slots[mds[sIn.mdIn].slot].info.subdet == SUB_BARREL;
        bool b1 = slots[mds[sIn.mdOut].slot].info.subdet == SUB_BARREL;
        bool b2 = slots[mds[sOut.mdIn].slot].info.subdet == SUB_BARREL;
        bool b3 = slots[mds[sOut.mdOut].slot].info.subdet == SUB_BARREL;
        if (rec) {  // the map's reset (Tracklet.cc:152-248) and "algo" (:256-293)
            float* tk = rec + kTLTyped;
            for (int k = 0; k < kTLKeys; ++k) tk[k] = -999.f;
            for (int k : kTLv2Keys) tk[k] = std::nanf("");  // absent unless _v2 runs
            tk[TK_algo] = (b0 && b1 && b2) ? 1.f : (b0 ? 2.f : 3.f);
        }
        if (b0 && b1 && b2) return tlBBBB(sIn, sOut, tr);          // BBBB, BBBE
        if (b0 && b1 && !b2 && !b3) return tlBBEE(sIn, sOut, tr);  // BBEE
        if (b0 && !b1 && !b2 && !b3) return tlBBEE(sIn, sOut, tr); // BEEE
        if (!b0 && !b1 && !b2 && !b3) return tlEEEE(sIn, sOut, tr);
        if (tr) tr->add(99, kNone, kNone, kNone, false);
        return false;                                              // other mixes fail (A-14)
    }

    // MiniDoublet::isIdxMatched (MiniDoublet.cc:686-693)
    bool mdMatch(int a, int b) const {
        return hits[mds[a].lo].idx == hits[mds[b].lo].idx && hits[mds[a].up].idx == hits[mds[b].up].idx;
    }
    // Segment::isIdxMatched (Segment.cc:868-875)
    bool sgMatch(int a, int b) const { return mdMatch(sgs[a].mdIn, sgs[b].mdIn) && mdMatch(sgs[a].mdOut, sgs[b].mdOut); }

    int lowerSlotOfMD(int md) const { return mds[md].slot; }

    // ---------------------------------------------------------------- stages
    // Event::createMiniDoublets (Event.cc:245-305)
    void createMiniDoublets() {
        std::vector<int> order = lowerOrder;
        for (int ls : order) {
            int us = getSlot(slots[ls].info.partner);
            const std::vector<int> loHits = slots[ls].hits;
            const std::vector<int> upHits = slots[us].hits;
            const ModInfo& m = slots[ls].info;
            for (int lo : loHits)
                for (int up : upHits) {
                    nCand[0]++;
                    countAt(0, binOfSlot(ls));
                    MDRec md;
                    if (acMD) {  // runMiniDoubletAllCombAlgo (MiniDoublet.cc:223-226): every candidate, values 0
                        md.lo = lo;
                        md.up = up;
                        md.slot = ls;
                        md.dz = md.shiftedDz = md.dphi = md.dphichange = 0;
                        // MiniDoublet::setAnchorHit (MiniDoublet.cc:60-81)
                        md.anchor = (m.type == TYPE_PS && m.layerType != LT_PIXEL) ? up : lo;
                        md.pseudo = true;
                    } else if (!mdAlgo(lo, up, md, nullptr)) {
                        continue;
                    }
                    countAt(5, binOfSlot(ls));
                    slots[ls].mds.push_back(int(mds.size()));
                    mds.push_back(md);
                }
        }
    }
    // Event::createPseudoMiniDoubletsFromAnchorModule (Event.cc:307-430): the
    // AllComb algorithm accepts every candidate (MiniDoublet.cc:223-226) and the
    // loops stop at the first, so each hit on the anchor side (the lower sensor,
    // or the upper one when the lower is the strip sensor of a PS module) pairs
    // with the first hit of the other sensor; the cached values stay 0
    // (MiniDoublet.cc:41-51).
    void createPseudoMiniDoublets() {
        std::vector<int> order = lowerOrder;
        for (int ls : order) {
            const ModInfo& m = slots[ls].info;
            const int us = getSlot(m.partner);
            const bool loopLower = !(m.type == TYPE_PS && m.layerType != LT_PIXEL);
            const std::vector<int> loHits = slots[ls].hits;
            const std::vector<int> upHits = slots[us].hits;
            const std::vector<int>& outer = loopLower ? loHits : upHits;
            const std::vector<int>& inner = loopLower ? upHits : loHits;
            for (int h : outer) {
                if (inner.empty()) break;
                nCand[0]++;
                countAt(0, binOfSlot(ls));
                countAt(5, binOfSlot(ls));
                MDRec md;
                md.lo = loopLower ? h : inner[0];
                md.up = loopLower ? inner[0] : h;
                md.slot = ls;
                md.dz = md.shiftedDz = md.dphi = md.dphichange = 0;
                md.anchor = loopLower ? md.lo : md.up;  // MiniDoublet::setAnchorHit (MiniDoublet.cc:60-81)
                md.pseudo = true;
                slots[ls].mds.push_back(int(mds.size()));
                mds.push_back(md);
            }
        }
    }
    // Event::createSegmentsFromInnerLowerModule (Event.cc:457-522)
    void createSegments() {
        for (int ls : lowerOrder) {
            const std::vector<uint32_t>& conn = g->connected(slots[ls].info.detId);
            for (int mi : slots[ls].mds)
                for (uint32_t d : conn) {
                    int os = findSlot(d);
                    if (os < 0) continue;
                    for (int mo : slots[os].mds) {
                        nCand[1]++;
                        countAt(1, binOfSlot(ls));
                        SGRec sg;
                        if (acSG ? allCombSG(mi, mo, sg) : sgAlgo(mi, mo, sg, nullptr)) {
                            countAt(6, binOfSlot(ls));
                            slots[ls].sgs.push_back(int(sgs.size()));
                            sgs.push_back(sg);
                        }
                    }
                }
        }
    }
    // Event::createTripletsFromInnerLowerModule (Event.cc:539-597) + Triplet.cc:50-122
    void createTriplets() {
        for (int ls : lowerOrder) {
            const std::vector<uint32_t>& conn = g->connected(slots[ls].info.detId);
            for (int si : slots[ls].sgs)
                for (uint32_t d : conn) {
                    int os = findSlot(d);
                    if (os < 0) continue;
                    for (int so : slots[os].sgs) {
                        nCand[2]++;
                        countAt(2, binOfSlot(ls));
                        // runTripletAllCombAlgo (Triplet.cc:45-48) accepts every candidate
                        if (acTP || mdMatch(sgs[si].mdOut, sgs[so].mdIn)) {
                            countAt(7, binOfSlot(ls));
                            slots[ls].tps.push_back(int(tps.size()));
                            tps.push_back({si, so});
                        }
                    }
                }
        }
    }
    // Event::createTrackletsFromInnerLowerModule (Event.cc:627-693)
    void createTracklets() {
        for (int ls : lowerOrder) {
            for (int si : slots[ls].sgs) {
                uint32_t bId = slots[mds[sgs[si].mdOut].slot].info.detId;
                for (uint32_t d : g->connected(bId)) {
                    int os = findSlot(d);
                    if (os < 0) continue;
                    for (int so : slots[os].sgs) {
                        nCand[3]++;
                        countAt(3, binOfSlot(ls));
                        if (acTL || tlAlgo(sgs[si], sgs[so], nullptr)) {
                            countAt(8, binOfSlot(ls));
                            slots[ls].tls.push_back(int(tls.size()));
                            tls.push_back({si, so, acTL});
                            tlLayer.push_back(layerOfMD(sgs[si].mdIn));  // addTrackletToEvent (Event.cc:179-193)
                        }
                    }
                }
            }
        }
    }
    // Event::createTrackCandidatesFromInnerModulesFromTriplets (Event.cc:889-982)
    void createTCFromTriplets() {
        for (int ls : lowerOrder) {
            for (int ti : slots[ls].tps) {
                int mdC = sgs[tps[ti].out].mdOut;
                uint32_t cId = slots[mds[mdC].slot].info.detId;
                for (uint32_t d : g->connected(cId)) {
                    int os = findSlot(d);
                    if (os < 0) continue;
                    for (int to : slots[os].tps) {
                        nCand[4]++;
                        countAt(4, binOfSlot(ls));
                        SGRec bridge;
                        if (!sgAlgo(mdC, sgs[tps[to].in].mdIn, bridge, nullptr)) continue;
                        if (!tlAlgo(bridge, sgs[tps[to].out], nullptr)) continue;
                        countAt(9, binOfSlot(ls));
                        tcs.push_back({0, {ti, to}});
                    }
                }
            }
        }
    }
    // Event::createTrackCandidatesFromInnerModulesFromTracklets (Event.cc:999-1062)
    // + TrackCandidate::runTrackCandidateDefaultAlgo (TrackCandidate.cc:143-226, A-13)
    void createTCFromTracklets() {
        for (int ls : lowerOrder) {
            for (int li : slots[ls].tls) {
                uint32_t bId = slots[mds[sgs[tls[li].in].mdOut].slot].info.detId;
                for (uint32_t d : g->connected(bId)) {
                    int os = findSlot(d);
                    if (os < 0) continue;
                    for (int lo : slots[os].tls) {
                        nCand[4]++;
                        countAt(4, binOfSlot(ls));
                        // runTrackCandidateAllCombAlgo (TrackCandidate.cc:138-141) accepts every candidate
                        if (acTC || sgMatch(tls[li].out, tls[lo].in)) {
                            countAt(9, binOfSlot(ls));
                            tcs.push_back({1, {li, lo, acTC}});
                        }
                    }
                }
            }
        }
    }

    // ---------------------------------------------------------- layer pairs
    // Layer lists (Layer.h): objects are appended to the layer of their inner
    // lower module as they are created (Event.cc:141-177); index 0-5 = barrel
    // layers 1-6, 6-10 = endcap layers 1-5.
    static int layerKey(int subdet, int layer) { return subdet == SUB_BARREL ? layer - 1 : 6 + layer - 1; }
    int layerOfMD(int md) const { const ModInfo& m = slots[mds[md].slot].info; return layerKey(m.subdet, m.layer); }
    std::vector<int> layerList(int key, int kind) const {  // kind 0: MDs, 1: SGs (creation order)
        std::vector<int> out;
        if (kind == 0) {
            for (size_t i = 0; i < mds.size(); ++i)
                if (layerOfMD(int(i)) == key) out.push_back(int(i));
        } else {
            for (size_t i = 0; i < sgs.size(); ++i)
                if (layerOfMD(sgs[i].mdIn) == key) out.push_back(int(i));
        }
        return out;
    }
    // SDL::Layer::segment_compatible_layer_pairs_ (Layer.cc:35-52), as (inner, outer) keys
    static const std::vector<std::pair<int, int>>& segmentLayerPairs() {
        static const std::vector<std::pair<int, int>> v = {
            {0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 5},        // barrel 1-2 ... 5-6
            {0, 6}, {1, 6}, {2, 6}, {3, 6}, {4, 6},        // barrel 1..5 - endcap 1
            {6, 7}, {7, 8}, {7, 9}, {7, 10}, {8, 9}, {9, 10}};  // endcap 1-2, 2-3, 2-4, 2-5, 3-4, 4-5
        return v;
    }
    // SDL::Layer::tracklet_compatible_layer_pairs_ (Layer.cc:3-33): barrel 1-3, 2-4, 3-5
    static const std::vector<std::pair<int, int>>& trackletLayerPairs() {
        static const std::vector<std::pair<int, int>> v = {{0, 2}, {1, 3}, {2, 4}};
        return v;
    }
    // Event::createSegments + createSegmentsFromTwoLayers (Event.cc:432-443, 793-821):
    // every MD of the inner layer against every MD of the outer layer; no counters.
    void createSegmentsLayers() {
        for (auto& pr : segmentLayerPairs()) {
            const std::vector<int> in = layerList(pr.first, 0), out = layerList(pr.second, 0);
            for (int mi : in)
                for (int mo : out) {
                    SGRec sg;
                    if (acSG ? allCombSG(mi, mo, sg) : sgAlgo(mi, mo, sg, nullptr)) {
                        slots[mds[mi].slot].sgs.push_back(int(sgs.size()));
                        sgs.push_back(sg);
                    }
                }
        }
    }
    // Event::createTracklets + createTrackletsFromTwoLayers (Event.cc:599-609, 755-790):
    // every SG of the inner layer against every SG of the outer layer; counted
    // on the inner layer; the TLs go to the inner layer's list only.
    std::vector<int> tlLayer;  // layer list of each TL (-1: none)
    void createTrackletsLayers() {
        for (auto& pr : trackletLayerPairs()) {
            const std::vector<int> in = layerList(pr.first, 1), out = layerList(pr.second, 1);
            for (int si : in)
                for (int so : out) {
                    nCand[3]++;
                    countAt(3, pr.first);
                    if (acTL || tlAlgo(sgs[si], sgs[so], nullptr)) {
                        countAt(8, pr.first);
                        tls.push_back({si, so, acTL});
                        tlLayer.push_back(pr.first);
                    }
                }
        }
    }
    // Event::createTrackCandidates + createTrackCandidatesFromTwoLayers (Event.cc:823-872):
    // the TLs of barrel layer 1 against those of barrel layer 3 with the default
    // algorithm, which passes iff the inner TL's outer SG is the outer TL's inner SG
    // (TrackCandidate.cc:143-226, A-13).
    void createTrackCandidatesLayers() {
        std::vector<int> in, out;
        for (size_t i = 0; i < tls.size(); ++i) {
            if (tlLayer[i] == 0) in.push_back(int(i));
            if (tlLayer[i] == 2) out.push_back(int(i));
        }
        for (int li : in)
            for (int lo : out) {
                nCand[4]++;
                countAt(4, 0);  // the inner layer, barrel 1
                if (acTC || sgMatch(tls[li].out, tls[lo].in)) {
                    countAt(9, 0);
                    tcs.push_back({1, {li, lo, acTC}});
                }
            }
    }
    // Event::createTrackletsViaNavigation (Event.cc:696-751): inner SG s1, every
    // SG s2 starting at s1's outer MD, every SG s3 starting at s2's outer MD
    // (the MDs' outward segment lists, filled in creation order by
    // Segment::addSelfPtrToMiniDoublets); the tracklet is (s1, s3), counted and
    // stored at s1's inner lower module.
    void createTrackletsNavigation() {
        std::vector<std::vector<int>> outward(mds.size());
        for (size_t k = 0; k < sgs.size(); ++k) outward[sgs[k].mdIn].push_back(int(k));
        for (int ls : lowerOrder)
            for (int si : slots[ls].sgs)
                for (int s2 : outward[sgs[si].mdOut])
                    for (int s3 : outward[sgs[s2].mdOut]) {
                        nCand[3]++;
                        countAt(3, binOfSlot(ls));
                        if (acTL || tlAlgo(sgs[si], sgs[s3], nullptr)) {
                            countAt(8, binOfSlot(ls));
                            slots[ls].tls.push_back(int(tls.size()));
                            tls.push_back({si, s3, acTL});
                            tlLayer.push_back(layerOfMD(sgs[si].mdIn));
                        }
                    }
    }

    void runLayers() {
        createMiniDoublets();
        createSegmentsLayers();
        createTrackletsLayers();
        createTrackCandidatesLayers();
    }
    // Stages in the given order: 0-5 the module-map stages, 6-8 createSegments,
    // createTracklets, createTrackCandidates, 9 createTrackletsViaNavigation,
    // 10 createPseudoMiniDoubletsFromAnchorModule; stage + 100 * algo runs the
    // stage with that Algo.h value (1: AllComb, for the tracklet stages
    // 1: DefaultNm1 and 2: AllComb; createTrackCandidatesFromTriplets and the
    // pseudo mini-doublets take no algorithm, Event.cc:874-982, 307-430).
    void runSequence(const int* seq, int n) {
        for (int k = 0; k < n; ++k) {
            if (seq[k] >= 100) {
                const int stage = seq[k] % 100, algo = seq[k] / 100;
                const bool tl = stage == 3 || stage == 7 || stage == 9;
                acMD = (stage == 0) && algo == 1;
                acSG = (stage == 1 || stage == 6) && algo == 1;
                acTP = stage == 2 && algo == 1;
                acTL = tl && algo == 2;
                acTC = (stage == 5 || stage == 8) && algo == 1;
                tlNm1 = tl && algo == 1;
                runSequence(&stage, 1);
                acMD = acSG = acTP = acTL = acTC = tlNm1 = false;
                continue;
            }
            switch (seq[k]) {
                case 0: createMiniDoublets(); break;
                case 1: createSegments(); break;
                case 2: createTriplets(); break;
                case 3: createTracklets(); break;
                case 4: createTCFromTriplets(); break;
                case 5: createTCFromTracklets(); break;
                case 6: createSegmentsLayers(); break;
                case 7: createTrackletsLayers(); break;
                case 8: createTrackCandidatesLayers(); break;
                case 9: createTrackletsNavigation(); break;
                case 10: createPseudoMiniDoublets(); break;
                // the three tracklet stages with DefaultNm1_TLAlgo
                case 11: tlNm1 = true; createTracklets(); tlNm1 = false; break;
                case 12: tlNm1 = true; createTrackletsLayers(); tlNm1 = false; break;
                case 13: tlNm1 = true; createTrackletsNavigation(); tlNm1 = false; break;
                default: break;
            }
        }
    }

    void run(int stages) {
        if (stages >= 1) createMiniDoublets();
        if (stages >= 2) createSegments();
        if (stages >= 3) createTriplets();
        if (stages >= 4) createTracklets();
        if (stages >= 5) createTCFromTriplets();
        if (stages >= 6) createTCFromTracklets();
    }

    // ------------------------------------------------------------- exporting
    void putMD(std::vector<int>& o, int m) const { o.push_back(hits[mds[m].lo].idx); o.push_back(hits[mds[m].up].idx); }
    void putSG(std::vector<int>& o, int s) const { putMD(o, sgs[s].mdIn); putMD(o, sgs[s].mdOut); }
    // Per-object variables of the stored MDs, SGs or TLs (the typed getters and
    // reco variables of MiniDoublet.h, Segment.h, Tracklet.h as the reference's
    // default algorithms leave them on a stored object), kMDVars / kSGVars /
    // kTLVars floats per object in object order: each object is evaluated once
    // more with the recording hooks on.  Pseudo MDs keep their zero typed
    // values and have no miniCut (NaN).
    std::vector<float> objectVars(int stage) {
        std::vector<float> out;
        const int nv = stage == ORACLE_MD ? kMDVars : stage == ORACLE_SG ? kSGVars : stage == ORACLE_TL ? kTLVars
                       : stage == ORACLE_TCL ? kTCLVars : 0;
        if (nv == 0) return out;
        float buf[kTLVars];
        rec = buf;
        auto put = [&] { out.insert(out.end(), buf, buf + nv); };
        if (stage == ORACLE_MD)
            for (const MDRec& md : mds) {
                std::fill(buf, buf + kTLVars, 0.f);
                if (md.pseudo) {
                    buf[6] = std::nanf("");
                } else {
                    MDRec t;
                    mdAlgo(md.lo, md.up, t, nullptr);
                }
                put();
            }
        if (stage == ORACLE_SG)
            for (const SGRec& sg : sgs) {
                std::fill(buf, buf + kTLVars, 0.f);
                if (sg.allComb) {  // constructor values, no reco variable
                    for (int k = 14; k < kSGVars; ++k) buf[k] = std::nanf("");
                } else {
                    SGRec t;
                    sgAlgo(sg.mdIn, sg.mdOut, t, nullptr);
                }
                put();
            }
        if (stage == ORACLE_TL) {
            const bool nm1 = tlNm1;
            tlNm1 = true;  // every stored TL passed; under DefaultNm1 it may fail the waived cut
            for (const PairRec& tl : tls) {
                std::fill(buf, buf + kTLVars, 0.f);
                if (tl.allComb)  // constructor values, an empty reco-variable map
                    for (int k = kTLTyped; k < kTLVars; ++k) buf[k] = std::nanf("");
                else
                    tlAlgo(sgs[tl.in], sgs[tl.out], nullptr);
                put();
            }
            tlNm1 = nm1;
        }
        if (stage == ORACLE_TCL)
            for (const auto& tc : tcs)
                if (tc.first == 1) {
                    if (tc.second.allComb)
                        buf[0] = buf[1] = std::nanf("");
                    else
                        tcRecoVars(tls[tc.second.in], tls[tc.second.out], buf);
                    put();
                }
        rec = nullptr;
        return out;
    }

    std::vector<int> tuples(int stage) const {
        std::vector<int> o;
        if (stage == ORACLE_MD) for (size_t i = 0; i < mds.size(); ++i) putMD(o, int(i));
        if (stage == ORACLE_SG) for (size_t i = 0; i < sgs.size(); ++i) putSG(o, int(i));
        if (stage == ORACLE_TP) for (auto& t : tps) { putSG(o, t.in); putSG(o, t.out); }
        if (stage == ORACLE_TL) for (auto& t : tls) { putSG(o, t.in); putSG(o, t.out); }
        for (auto& tc : tcs) {
            if (stage == ORACLE_TCP && tc.first == 0) {
                putSG(o, tps[tc.second.in].in); putSG(o, tps[tc.second.in].out);
                putSG(o, tps[tc.second.out].in); putSG(o, tps[tc.second.out].out);
            }
            if (stage == ORACLE_TCL && tc.first == 1) {
                putSG(o, tls[tc.second.in].in); putSG(o, tls[tc.second.in].out);
                putSG(o, tls[tc.second.out].in); putSG(o, tls[tc.second.out].out);
            }
        }
        return o;
    }
};

struct Result {
    std::vector<int> t[6];
    std::vector<float> vars[6];  // per-object variables of MDs, SGs, TLs and TCLs (index = stage)
    unsigned long long counters[10];
    unsigned long long layers[110];  // Event::layerCnt
};

void fillEvent(Event& ev, const Geometry* g, int libm, int n, const float* x, const float* y, const float* z,
               const unsigned* detId, const int* idx) {
    ev.g = g;
    ev.lm.mode = libm;
    ev.hits.reserve(n);
    for (int i = 0; i < n; ++i) ev.addHit(x[i], y[i], z[i], idx[i], detId[i]);
}

}  // namespace oracle

// The op-counting build (opcount.cc) includes this file with float / double
// replaced by counting types and provides its own entry point instead.
#ifndef ORACLE_OPCOUNT
using namespace oracle;

extern "C" {

// Known-answer hooks of single functions of the restatement (closed-form and
// compiled-reference checks of the tests; oracle/kat_check.cc).
// MathUtil::Phi_mpi_pi, MathUtil::ATan2 (MathUtil.cc:3-27)
float oracle_kat_phi_mpi_pi(float x) { return phiMpiPi(x); }
float oracle_kat_atan2(float y, float x) { return atan2Sdl(y, x); }
// MathUtil::getCenterFromThreePoints (MathUtil.cc:44-106) on the points (x[k], y[k]), k = 0..2
void oracle_kat_center(const float* x, const float* y, float* out2) {
    Event::centerFromThree(P3{x[0], y[0], 0.f}, P3{x[1], y[1], 0.f}, P3{x[2], y[2], 0.f}, out2[0], out2[1]);
}
// MiniDoublet::moduleGapSize and isTighterTiltedModules (MiniDoublet.cc:1242-1332) of a module
float oracle_kat_module_gap(unsigned detId, int* tighter) {
    const ModInfo m = decode(detId);
    if (tighter) *tighter = isTighterTilted(m) ? 1 : 0;
    return moduleGapSize(m);
}
// MiniDoublet::shiftStripHits (MiniDoublet.cc:989-1209) of a lower-module hit
// and an upper-module hit with the given geometry (glibc libm): out = x, y, z
void oracle_kat_shift(void* geom, unsigned lowerDetId, const float* lo, const float* up, float* out3) {
    Event ev;
    ev.g = static_cast<Geometry*>(geom);
    ev.lm.mode = ORACLE_LIBM_GLIBC;
    const ModInfo m = decode(lowerDetId);
    ev.addHit(lo[0], lo[1], lo[2], 0, lowerDetId);
    ev.addHit(up[0], up[1], up[2], 1, m.partner);
    ev.shiftStripHits(ev.hits[0], ev.hits[1], ev.slots[ev.hits[0].slot], out3[0], out3[1], out3[2]);
}

// The restated detId decoding, fields in the order subdet, side, layer, rod,
// ring, module, isLower, isInverted, partnerDetId, moduleType, moduleLayerType.
void oracle_module_decode(unsigned d, unsigned out[11]) {
    const ModInfo m = decode(d);
    const unsigned v[11] = {unsigned(m.subdet), unsigned(m.side), unsigned(m.layer), unsigned(m.rod),
                            unsigned(m.ring), unsigned(m.module), unsigned(m.isLower), unsigned(m.isInverted),
                            m.partner, unsigned(m.type), unsigned(m.layerType)};
    for (int k = 0; k < 11; ++k) out[k] = v[k];
}

void* oracle_geom_load(const char* tiltedPath, const char* endcapPath, const char* mapPath) {
    auto* g = new Geometry();
    if (!loadGeometry(*g, tiltedPath, endcapPath, mapPath)) { delete g; return nullptr; }
    return g;
}

void oracle_geom_free(void* geom) { delete static_cast<Geometry*>(geom); }

void* oracle_run(void* geom, int nhits, const float* x, const float* y, const float* z, const unsigned* detId,
                 const int* idx, int libm, int stages) {
    Event ev;
    fillEvent(ev, static_cast<Geometry*>(geom), libm, nhits, x, y, z, detId, idx);
    if (stages == ORACLE_LAYER_CHAIN) ev.runLayers();
    else ev.run(stages);
    auto* r = new Result();
    for (int s = 0; s < 6; ++s) r->t[s] = ev.tuples(s);
    for (int s : {ORACLE_MD, ORACLE_SG, ORACLE_TL, ORACLE_TCL}) r->vars[s] = ev.objectVars(s);
    for (int s = 0; s < 5; ++s) r->counters[s] = ev.nCand[s];
    r->counters[5] = ev.mds.size();
    r->counters[6] = ev.sgs.size();
    r->counters[7] = ev.tps.size();
    r->counters[8] = ev.tls.size();
    r->counters[9] = ev.tcs.size();
    std::memcpy(r->layers, ev.layerCnt, sizeof(r->layers));
    return r;
}

void* oracle_run_sequence(void* geom, int nhits, const float* x, const float* y, const float* z,
                          const unsigned* detId, const int* idx, int libm, const int* seq, int n) {
    Event ev;
    fillEvent(ev, static_cast<Geometry*>(geom), libm, nhits, x, y, z, detId, idx);
    ev.runSequence(seq, n);
    auto* r = new Result();
    for (int s = 0; s < 6; ++s) r->t[s] = ev.tuples(s);
    for (int s : {ORACLE_MD, ORACLE_SG, ORACLE_TL, ORACLE_TCL}) r->vars[s] = ev.objectVars(s);
    for (int s = 0; s < 5; ++s) r->counters[s] = ev.nCand[s];
    r->counters[5] = ev.mds.size();
    r->counters[6] = ev.sgs.size();
    r->counters[7] = ev.tps.size();
    r->counters[8] = ev.tls.size();
    r->counters[9] = ev.tcs.size();
    std::memcpy(r->layers, ev.layerCnt, sizeof(r->layers));
    return r;
}

static const int kWidth[6] = {2, 4, 8, 8, 16, 16};

int oracle_var_count(int stage) {
    return stage == ORACLE_MD ? kMDVars : stage == ORACLE_SG ? kSGVars : stage == ORACLE_TL ? kTLVars
           : stage == ORACLE_TCL ? kTCLVars : 0;
}
void oracle_result_vars(void* res, int stage, float* out) {
    if (stage < 0 || stage > ORACLE_TCL) return;
    const std::vector<float>& v = static_cast<Result*>(res)->vars[stage];
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(float));
}

long oracle_result_count(void* res, int stage) {
    return long(static_cast<Result*>(res)->t[stage].size() / kWidth[stage]);
}
void oracle_result_copy(void* res, int stage, int* out) {
    const std::vector<int>& v = static_cast<Result*>(res)->t[stage];
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(int));
}
void oracle_result_counters(void* res, unsigned long long* out10) {
    std::memcpy(out10, static_cast<Result*>(res)->counters, sizeof(unsigned long long) * 10);
}
void oracle_result_layers(void* res, unsigned long long* out110) {
    std::memcpy(out110, static_cast<Result*>(res)->layers, sizeof(unsigned long long) * 110);
}
void oracle_result_free(void* res) { delete static_cast<Result*>(res); }

int oracle_explain(void* geom, int nhits, const float* x, const float* y, const float* z, const unsigned* detId,
                   const int* idx, int kind, const int* a, const int* b, int libmB, double* out) {
    Geometry* g = static_cast<Geometry*>(geom);
    Trace trs[2];
    int decided[2] = {-1, -1};
    for (int mode = 0; mode < 2; ++mode) {
        Event ev;
        fillEvent(ev, g, mode == 0 ? ORACLE_LIBM_GLIBC : libmB, nhits, x, y, z, detId, idx);
        std::unordered_map<int, int> pos;  // user idx -> hit position
        for (int i = 0; i < nhits; ++i) pos[idx[i]] = i;
        auto makeMD = [&](int lo, int up) -> int {
            MDRec md;
            ev.mdAlgo(pos.at(lo), pos.at(up), md, nullptr);
            ev.mds.push_back(md);
            return int(ev.mds.size()) - 1;
        };
        bool ok;
        if (kind == 0) {
            MDRec md;
            ok = ev.mdAlgo(pos.at(a[0]), pos.at(a[1]), md, &trs[mode]);
        } else if (kind == 1) {
            int m0 = makeMD(a[0], a[1]), m1 = makeMD(a[2], a[3]);
            SGRec sg;
            ok = ev.sgAlgo(m0, m1, sg, &trs[mode]);
        } else {
            int m0 = makeMD(a[0], a[1]), m1 = makeMD(a[2], a[3]), m2 = makeMD(b[0], b[1]), m3 = makeMD(b[2], b[3]);
            SGRec s0, s1;
            ev.sgAlgo(m0, m1, s0, nullptr);
            ev.sgAlgo(m2, m3, s1, nullptr);
            ok = ev.tlAlgo(s0, s1, &trs[mode]);
        }
        decided[mode] = ok ? -1 : trs[mode].failed;
    }
    // the deciding cut: the one that failed under one policy
    int cut = decided[0] >= 0 ? decided[0] : decided[1];
    out[0] = decided[0];
    out[1] = decided[1];
    for (int k = 2; k < 9; ++k) out[k] = kNone;
    if (cut < 0) return 1;
    for (int mode = 0; mode < 2; ++mode)
        for (int i = 0; i < trs[mode].n; ++i)
            if (trs[mode].id[i] == cut) {
                out[2 + 3 * mode] = trs[mode].v[i];
                out[3 + 3 * mode] = trs[mode].lo[i];
                out[4 + 3 * mode] = trs[mode].hi[i];
            }
    // distance (in float ulps at the threshold) between variable and nearest bound, GLIBC policy
    double v = out[2], best = 1e300;
    for (int k = 3; k <= 4; ++k) {
        double t = out[k];
        if (std::isnan(t) || std::isnan(v)) continue;
        float tf = float(t);
        float ulp = std::nextafter(std::abs(tf), INFINITY) - std::abs(tf);
        best = std::min(best, std::abs(v - t) / double(ulp));
    }
    out[8] = best;
    return 0;
}

double oracle_time(void* geom, int nevents, const long* offsets, const float* x, const float* y, const float* z,
                   const unsigned* detId, const int* idx, int libm, int threads) {
    Geometry* g = static_cast<Geometry*>(geom);
    std::atomic<int> next(0);
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&]() {
        for (;;) {
            int e = next.fetch_add(1);
            if (e >= nevents) break;
            long o = offsets[e], n = offsets[e + 1] - offsets[e];
            Event ev;
            fillEvent(ev, g, libm, int(n), x + o, y + o, z + o, detId + o, idx + o);
            ev.run(6);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < std::max(1, threads); ++t) th.emplace_back(work);
    for (auto& t : th) t.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
#endif  // ORACLE_OPCOUNT
