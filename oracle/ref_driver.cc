// Harness around the compiled, unmodified reference (oracle/_ref/libsdl_ref.so).
//
// It is test infrastructure: it generates the same seeded events as the rest of
// this repository (sdl_amd/csrc/gen), feeds them hit by hit through the
// reference's own API (SDL::Event::addHitToModule, Event.cc:120-139) after
// loading the three geometry singletons from the generated text files
// (TiltedGeometry.cc:18-41, EndcapGeometry.cc:18-61, ModuleConnectionMap.cc:18-47),
// runs the module-map chain
//     createMiniDoublets, createSegmentsWithModuleMap, createTriplets,
//     createTrackletsWithModuleMap, createTrackCandidatesFromTriplets,
//     createTrackCandidatesFromTracklets                       (Event.h:240-300)
// and writes every object as a tuple of hit indices, plus the per-stage counters.
// Mode dumplayers runs the layer-pair chain instead
//     createMiniDoublets, createSegments, createTracklets, createTrackCandidates
//                                      (Event.cc:245-255, 432-443, 599-609, 823-837)
// and adds the per-layer counters of every stage.
//
// Usage:
//   sdl_ref_driver dump        <workload> <geomdir> <event0> <nevents> <outfile>
//   sdl_ref_driver dumplayers  <workload> <geomdir> <event0> <nevents> <outfile>
//   sdl_ref_driver dumpnav     <workload> <geomdir> <event0> <nevents> <outfile>
//   sdl_ref_driver dumppseudo  <workload> <geomdir> <event0> <nevents> <outfile>
//     (createPseudoMiniDoubletsFromAnchorModule, then the module-map chain)
//     (createMiniDoublets, createSegmentsWithModuleMap, createTrackletsViaNavigation)
//   sdl_ref_driver dumpnm1     <workload> <geomdir> <event0> <nevents> <outfile>
//     (the module-map chain with createTrackletsWithModuleMap(DefaultNm1_TLAlgo))
//   sdl_ref_driver dumpvars    <workload> <geomdir> <event0> <nevents> <outfile>
//     (the module-map chain; every MD, SG and TL with its typed per-object
//     values and reco variables, TLs with their whole reco-variable map, TCs
//     from tracklets with theirs, as float bit patterns)
//   sdl_ref_driver dumpseq     <workload> <geomdir> <event0> <nevents> <outfile> <sequence>
//     (the create* calls of <sequence>, comma-separated stage:algo tokens with
//     the Algo.h enum values; stages md, pmd, sg, sgl, tp, tl, tll, tln, tcp,
//     tcl, tcll for createMiniDoublets, createPseudoMiniDoubletsFromAnchorModule,
//     createSegmentsWithModuleMap, createSegments, createTriplets,
//     createTrackletsWithModuleMap, createTracklets, createTrackletsViaNavigation,
//     createTrackCandidatesFromTriplets, createTrackCandidatesFromTracklets,
//     createTrackCandidates; objects, counters and per-layer counters)
//   sdl_ref_driver time        <workload> <geomdir> <event0> <nevents> <threads> [stages]
// `stages` limits the chain to its first k stages (1..6, default 6).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>
#include <atomic>
#include <cmath>
#include <map>

#include "Event.h"
#include "EndcapGeometry.h"
#include "ModuleConnectionMap.h"
#include "TiltedGeometry.h"
#include "Tracklet.h"
#include "TrackCandidate.h"
#include "sdl_gen.h"

namespace {

bool fileExists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

int geomKindOf(int workload) { return sdlgen::geometryKindOf(workload); }

void loadGeometry(int workload, const std::string& dir, sdlgen::Geometry& g) {
    int kind = geomKindOf(workload);
    sdlgen::buildGeometry(kind, g, kind == 1 ? 50000 : 200000);
    std::string t = dir + "/tilted_" + std::to_string(kind) + ".txt";
    std::string e = dir + "/endcap_" + std::to_string(kind) + ".txt";
    std::string m = dir + "/map_" + std::to_string(kind) + ".txt";
    sdlgen::writeGeometryFiles(g, t, e, m);
    SDL::tiltedGeometry.load(t);
    SDL::endcapGeometry.load(e);
    SDL::moduleConnectionMap.load(m);
    // Touch every key once so that later lookups (std::map::operator[]) never
    // insert; this makes the singletons safe to read from several threads.
    for (const sdlgen::Sensor& s : g.sensors) {
        SDL::tiltedGeometry.getDrDz(s.detId);
        SDL::tiltedGeometry.getSlope(s.detId);
        SDL::endcapGeometry.getSlopeLower(s.detId);
        SDL::endcapGeometry.getCentroidPhi(s.detId);
        SDL::moduleConnectionMap.getConnectedModuleDetIds(s.detId);
    }
}

void runChain(SDL::Event& ev, int stages) {
    if (stages >= 1) ev.createMiniDoublets();
    if (stages >= 2) ev.createSegmentsWithModuleMap();
    if (stages >= 3) ev.createTriplets();
    if (stages >= 4) ev.createTrackletsWithModuleMap();
    if (stages >= 5) ev.createTrackCandidatesFromTriplets();
    if (stages >= 6) ev.createTrackCandidatesFromTracklets();
}

void fill(SDL::Event& sev, const sdlgen::Event& ev) {
    for (size_t i = 0; i < ev.x.size(); ++i)
        sev.addHitToModule(SDL::Hit(ev.x[i], ev.y[i], ev.z[i], ev.idx[i]), ev.detId[i]);
}

void putMD(FILE* f, const SDL::MiniDoublet* md) {
    std::fprintf(f, " %d %d", md->lowerHitPtr()->idx(), md->upperHitPtr()->idx());
}
void putSG(FILE* f, const SDL::Segment* sg) {
    putMD(f, sg->innerMiniDoubletPtr());
    putMD(f, sg->outerMiniDoubletPtr());
}

void runLayerChain(SDL::Event& ev) {
    ev.createMiniDoublets();
    ev.createSegments();
    ev.createTracklets();
    ev.createTrackCandidates();
}

// Per-layer counters (barrel 1-6, endcap 1-5) of the ten counter kinds.
void dumpLayerCounters(FILE* f, SDL::Event& sev) {
    typedef unsigned (SDL::Event::*Getter)(unsigned);
    const Getter barrel[10] = {&SDL::Event::getNumberOfMiniDoubletCandidatesByLayerBarrel,
                               &SDL::Event::getNumberOfSegmentCandidatesByLayerBarrel,
                               &SDL::Event::getNumberOfTripletCandidatesByLayerBarrel,
                               &SDL::Event::getNumberOfTrackletCandidatesByLayerBarrel,
                               &SDL::Event::getNumberOfTrackCandidateCandidatesByLayerBarrel,
                               &SDL::Event::getNumberOfMiniDoubletsByLayerBarrel,
                               &SDL::Event::getNumberOfSegmentsByLayerBarrel,
                               &SDL::Event::getNumberOfTripletsByLayerBarrel,
                               &SDL::Event::getNumberOfTrackletsByLayerBarrel,
                               &SDL::Event::getNumberOfTrackCandidatesByLayerBarrel};
    const Getter endcap[10] = {&SDL::Event::getNumberOfMiniDoubletCandidatesByLayerEndcap,
                               &SDL::Event::getNumberOfSegmentCandidatesByLayerEndcap,
                               &SDL::Event::getNumberOfTripletCandidatesByLayerEndcap,
                               &SDL::Event::getNumberOfTrackletCandidatesByLayerEndcap,
                               &SDL::Event::getNumberOfTrackCandidateCandidatesByLayerEndcap,
                               &SDL::Event::getNumberOfMiniDoubletsByLayerEndcap,
                               &SDL::Event::getNumberOfSegmentsByLayerEndcap,
                               &SDL::Event::getNumberOfTripletsByLayerEndcap,
                               &SDL::Event::getNumberOfTrackletsByLayerEndcap,
                               &SDL::Event::getNumberOfTrackCandidatesByLayerEndcap};
    std::fprintf(f, "L");
    for (int k = 0; k < 10; ++k) {
        for (unsigned i = 0; i < 6; ++i) std::fprintf(f, " %u", (sev.*barrel[k])(i));
        for (unsigned i = 0; i < 5; ++i) std::fprintf(f, " %u", (sev.*endcap[k])(i));
    }
    std::fprintf(f, "\n");
}

// Layer-pair chain: MDs and SGs through the modules (addSegmentToEvent adds to
// both, Event.cc:163-177), TLs and TCs through the layers (addTrackletToLowerLayer,
// addTrackCandidateToLowerLayer, Event.cc:227-243), then the per-layer counters
// (barrel 1-6, endcap 1-5) of the ten counter kinds.
void dumpLayerEvent(FILE* f, SDL::Event& sev, long ievent) {
    std::fprintf(f, "E %ld\n", ievent);
    for (SDL::Module* m : sev.getModulePtrs()) {
        for (SDL::MiniDoublet* md : m->getMiniDoubletPtrs()) { std::fprintf(f, "MD"); putMD(f, md); std::fprintf(f, "\n"); }
        for (SDL::Segment* sg : m->getSegmentPtrs()) { std::fprintf(f, "SG"); putSG(f, sg); std::fprintf(f, "\n"); }
    }
    for (SDL::Layer* L : sev.getLayerPtrs()) {
        for (SDL::Tracklet* tl : L->getTrackletPtrs()) {
            std::fprintf(f, "TL"); putSG(f, tl->innerSegmentPtr()); putSG(f, tl->outerSegmentPtr()); std::fprintf(f, "\n");
        }
        for (SDL::TrackCandidate* tc : L->getTrackCandidatePtrs()) {
            SDL::TrackletBase* a = tc->innerTrackletBasePtr();
            SDL::TrackletBase* b = tc->outerTrackletBasePtr();
            std::fprintf(f, "TCL");
            putSG(f, a->innerSegmentPtr()); putSG(f, a->outerSegmentPtr());
            putSG(f, b->innerSegmentPtr()); putSG(f, b->outerSegmentPtr());
            std::fprintf(f, "\n");
        }
    }
    std::fprintf(f, "N %u %u %u %u %u\n", sev.getNumberOfMiniDoubletCandidates(), sev.getNumberOfSegmentCandidates(),
                 sev.getNumberOfTripletCandidates(), sev.getNumberOfTrackletCandidates(),
                 sev.getNumberOfTrackCandidateCandidates());
    std::fprintf(f, "C %u %u %u %u %u\n", sev.getNumberOfMiniDoublets(), sev.getNumberOfSegments(),
                 sev.getNumberOfTriplets(), sev.getNumberOfTracklets(), sev.getNumberOfTrackCandidates());
    dumpLayerCounters(f, sev);
}

// Any sequence of stages: MDs, SGs and TPs through the modules, TLs and TCs
// through the layers (every TL and TC is added to one, Event.cc:179-243).
void dumpSeqEvent(FILE* f, SDL::Event& sev, long ievent) {
    std::fprintf(f, "E %ld\n", ievent);
    for (SDL::Module* m : sev.getModulePtrs()) {
        for (SDL::MiniDoublet* md : m->getMiniDoubletPtrs()) { std::fprintf(f, "MD"); putMD(f, md); std::fprintf(f, "\n"); }
        for (SDL::Segment* sg : m->getSegmentPtrs()) { std::fprintf(f, "SG"); putSG(f, sg); std::fprintf(f, "\n"); }
        for (SDL::Triplet* tp : m->getTripletPtrs()) {
            std::fprintf(f, "TP"); putSG(f, tp->innerSegmentPtr()); putSG(f, tp->outerSegmentPtr()); std::fprintf(f, "\n");
        }
    }
    for (SDL::Layer* L : sev.getLayerPtrs()) {
        for (SDL::Tracklet* tl : L->getTrackletPtrs()) {
            std::fprintf(f, "TL"); putSG(f, tl->innerSegmentPtr()); putSG(f, tl->outerSegmentPtr()); std::fprintf(f, "\n");
        }
        for (SDL::TrackCandidate* tc : L->getTrackCandidatePtrs()) {
            SDL::TrackletBase* a = tc->innerTrackletBasePtr();
            SDL::TrackletBase* b = tc->outerTrackletBasePtr();
            std::fprintf(f, dynamic_cast<SDL::Triplet*>(a) != nullptr ? "TCP" : "TCL");
            putSG(f, a->innerSegmentPtr()); putSG(f, a->outerSegmentPtr());
            putSG(f, b->innerSegmentPtr()); putSG(f, b->outerSegmentPtr());
            std::fprintf(f, "\n");
        }
    }
    std::fprintf(f, "N %u %u %u %u %u\n", sev.getNumberOfMiniDoubletCandidates(), sev.getNumberOfSegmentCandidates(),
                 sev.getNumberOfTripletCandidates(), sev.getNumberOfTrackletCandidates(),
                 sev.getNumberOfTrackCandidateCandidates());
    std::fprintf(f, "C %u %u %u %u %u\n", sev.getNumberOfMiniDoublets(), sev.getNumberOfSegments(),
                 sev.getNumberOfTriplets(), sev.getNumberOfTracklets(), sev.getNumberOfTrackCandidates());
    dumpLayerCounters(f, sev);
}

// Runs one stage:algo token of a dumpseq sequence; false for an unknown stage.
bool runToken(SDL::Event& sev, const std::string& stage, int algo) {
    if (stage == "md") sev.createMiniDoublets(SDL::MDAlgo(algo));
    else if (stage == "pmd") sev.createPseudoMiniDoubletsFromAnchorModule(SDL::MDAlgo(algo));
    else if (stage == "sg") sev.createSegmentsWithModuleMap(SDL::SGAlgo(algo));
    else if (stage == "sgl") sev.createSegments(SDL::SGAlgo(algo));
    else if (stage == "tp") sev.createTriplets(SDL::TPAlgo(algo));
    else if (stage == "tl") sev.createTrackletsWithModuleMap(SDL::TLAlgo(algo));
    else if (stage == "tll") sev.createTracklets(SDL::TLAlgo(algo));
    else if (stage == "tln") sev.createTrackletsViaNavigation(SDL::TLAlgo(algo));
    else if (stage == "tcp") sev.createTrackCandidatesFromTriplets(SDL::TCAlgo(algo));
    else if (stage == "tcl") sev.createTrackCandidatesFromTracklets(SDL::TCAlgo(algo));
    else if (stage == "tcll") sev.createTrackCandidates(SDL::TCAlgo(algo));
    else return false;
    return true;
}

void dumpEvent(FILE* f, SDL::Event& sev, long ievent) {
    std::fprintf(f, "E %ld\n", ievent);
    // Objects are reached through the per-module pointer lists, which hold every
    // MD, SG, TP and TL (Event.cc:141-193); TCs through the layers (Event.cc:232-239).
    for (SDL::Module* m : sev.getModulePtrs()) {
        for (SDL::MiniDoublet* md : m->getMiniDoubletPtrs()) { std::fprintf(f, "MD"); putMD(f, md); std::fprintf(f, "\n"); }
        for (SDL::Segment* sg : m->getSegmentPtrs()) { std::fprintf(f, "SG"); putSG(f, sg); std::fprintf(f, "\n"); }
        for (SDL::Triplet* tp : m->getTripletPtrs()) {
            std::fprintf(f, "TP"); putSG(f, tp->innerSegmentPtr()); putSG(f, tp->outerSegmentPtr()); std::fprintf(f, "\n");
        }
        for (SDL::Tracklet* tl : m->getTrackletPtrs()) {
            std::fprintf(f, "TL"); putSG(f, tl->innerSegmentPtr()); putSG(f, tl->outerSegmentPtr()); std::fprintf(f, "\n");
        }
    }
    for (SDL::Layer* L : sev.getLayerPtrs()) {
        for (SDL::TrackCandidate* tc : L->getTrackCandidatePtrs()) {
            // A TC made from triplets stores two Triplets, one made from tracklets two
            // Tracklets (TrackCandidate.cc:38-88 throw on the wrong accessor).
            SDL::TrackletBase* a = tc->innerTrackletBasePtr();
            SDL::TrackletBase* b = tc->outerTrackletBasePtr();
            bool fromTP = dynamic_cast<SDL::Triplet*>(a) != nullptr;
            std::fprintf(f, fromTP ? "TCP" : "TCL");
            putSG(f, a->innerSegmentPtr()); putSG(f, a->outerSegmentPtr());
            putSG(f, b->innerSegmentPtr()); putSG(f, b->outerSegmentPtr());
            std::fprintf(f, "\n");
        }
    }
    std::fprintf(f, "N %u %u %u %u %u\n", sev.getNumberOfMiniDoubletCandidates(), sev.getNumberOfSegmentCandidates(),
                 sev.getNumberOfTripletCandidates(), sev.getNumberOfTrackletCandidates(),
                 sev.getNumberOfTrackCandidateCandidates());
    std::fprintf(f, "C %u %u %u %u %u\n", sev.getNumberOfMiniDoublets(), sev.getNumberOfSegments(),
                 sev.getNumberOfTriplets(), sev.getNumberOfTracklets(), sev.getNumberOfTrackCandidates());
}

// A reco variable of an object's map, NaN when the algorithm never set it.
float recoVar(const std::map<std::string, float>& m, const char* key) {
    auto it = m.find(key);
    return it == m.end() ? std::nanf("") : it->second;
}
void putBits(FILE* f, float v) {
    unsigned u;
    std::memcpy(&u, &v, 4);
    std::fprintf(f, " %08x", u);
}

// Per-object values of every MD, SG and TL (MiniDoublet.h, Segment.h,
// Tracklet.h getters and the setRecoVars maps), reached as in dumpEvent.
// The keys of a stored tracklet's reco-variable map (the union over the three
// default algorithms, in std::map order); a key a variant never sets is dumped
// as NaN, like an absent MD key.
const char* const kTrackletKeys[] = {
    "algo", "alpha1GeV_OutLo", "betaAv", "betaAv_0th", "betaAv_1st", "betaAv_2nd", "betaAv_3rd", "betaAv_4th",
    "betaIn", "betaInRHmax", "betaInRHmin", "betaIn_0th", "betaIn_1st", "betaIn_1stCorr", "betaIn_2nd",
    "betaIn_2ndCorr", "betaIn_3rd", "betaIn_3rdCorr", "betaIn_4th", "betaIn_cut", "betaOut", "betaOutRHmax",
    "betaOutRHmin", "betaOut_0th", "betaOut_1st", "betaOut_1stCorr", "betaOut_2nd", "betaOut_2ndCorr",
    "betaOut_3rd", "betaOut_3rdCorr", "betaOut_4th", "betaPt_0th", "betaPt_1st", "betaPt_2nd", "betaPt_3rd",
    "betaPt_4th", "betacormode", "dBeta", "dBetaCut2", "dBetaLum2", "dBetaMuls", "dBetaRIn2", "dBetaROut2",
    "dBetaRes", "dBeta_0th", "dBeta_1st", "dBeta_2nd", "dBeta_3rd", "dBeta_4th", "dBeta_midPoint", "dPhi",
    "deltaPhiPos", "deltaZLum", "dr", "drt_InSeg", "drt_tl_axis", "dzDrtScale", "hit1_x", "hit1_y", "hit2_x",
    "hit2_y", "hit3_x", "hit3_y", "hit4_x", "hit4_y", "innerSgInnerMdDetId", "innerSgOuterMdDetId",
    "k2Rinv1GeVf", "kRinv1GeVf", "kZ", "outerSgInnerMdDetId", "outerSgOuterMdDetId", "pixelPSZpitch", "ptCut",
    "pt_beta", "pt_betaIn", "pt_betaOut", "rawBetaIn", "rawBetaInCorrection", "rawBetaOut",
    "rawBetaOutCorrection", "rtHi_point", "rtIn", "rtLo_point", "rtOut", "rtOut_o_rtIn", "rtRatio_OutLoInLo",
    "rt_InLo", "rt_OutLo", "sdIn_alpha", "sdIn_d", "sdOut_alphaOut", "sdOut_d", "sdlCut", "sdlSlope",
    "sinAlphaMax", "strip2SZpitch", "zGeom", "zHi", "zHiPointed", "zIn", "zLo", "zLoPointed", "zOut", "z_InLo",
    "z_OutLo", "zpitch_InLo", "zpitch_OutLo"};

void dumpVars(FILE* f, SDL::Event& sev, long ievent) {
    std::fprintf(f, "E %ld\n", ievent);
    for (SDL::Module* m : sev.getModulePtrs()) {
        for (SDL::MiniDoublet* md : m->getMiniDoubletPtrs()) {
            std::fprintf(f, "MV");
            putMD(f, md);
            const float v[7] = {md->getDz(), md->getShiftedDz(), md->getDeltaPhi(), md->getDeltaPhiChange(),
                                md->getDeltaPhiNoShift(), md->getDeltaPhiChangeNoShift(),
                                recoVar(md->getRecoVars(), "miniCut")};
            for (float x : v) putBits(f, x);
            std::fprintf(f, "\n");
        }
        for (SDL::Segment* sg : m->getSegmentPtrs()) {
            std::fprintf(f, "SV");
            putSG(f, sg);
            const std::map<std::string, float>& r = sg->getRecoVars();
            const float v[22] = {sg->getRtIn(), sg->getRtOut(), sg->getZIn(), sg->getZOut(), sg->getZLo(), sg->getZHi(),
                                 sg->getRtLo(), sg->getRtHi(), sg->getDeltaPhi(), sg->getDeltaPhiMin(),
                                 sg->getDeltaPhiMax(), sg->getDeltaPhiChange(), sg->getDeltaPhiMinChange(),
                                 sg->getDeltaPhiMaxChange(), recoVar(r, "sdCut"), recoVar(r, "sdSlope"),
                                 recoVar(r, "sdMuls"), recoVar(r, "sdPVoff"), recoVar(r, "deltaPhi"),
                                 recoVar(r, "dAlpha_innerMD_segment"), recoVar(r, "dAlpha_outerMD_segment"),
                                 recoVar(r, "dAlpha_innerMD_outerMD")};
            for (float x : v) putBits(f, x);
            std::fprintf(f, "\n");
        }
        for (SDL::Tracklet* tl : m->getTrackletPtrs()) {
            std::fprintf(f, "TV");
            putSG(f, tl->innerSegmentPtr());
            putSG(f, tl->outerSegmentPtr());
            const float v[6] = {tl->getBetaIn(), tl->getBetaInCut(), tl->getBetaOut(), tl->getBetaOutCut(),
                                tl->getDeltaBeta(), tl->getDeltaBetaCut()};
            for (float x : v) putBits(f, x);
            for (const char* k : kTrackletKeys) putBits(f, recoVar(tl->getRecoVars(), k));
            std::fprintf(f, "\n");
        }
    }
    // TCs made from tracklets carry dR and innerR (TrackCandidate.cc:206-207);
    // those made from triplets run no algorithm and have an empty map
    for (SDL::Layer* L : sev.getLayerPtrs()) {
        for (SDL::TrackCandidate* tc : L->getTrackCandidatePtrs()) {
            SDL::TrackletBase* a = tc->innerTrackletBasePtr();
            SDL::TrackletBase* b = tc->outerTrackletBasePtr();
            if (dynamic_cast<SDL::Triplet*>(a) != nullptr) continue;
            std::fprintf(f, "CV");
            putSG(f, a->innerSegmentPtr()); putSG(f, a->outerSegmentPtr());
            putSG(f, b->innerSegmentPtr()); putSG(f, b->outerSegmentPtr());
            putBits(f, recoVar(tc->getRecoVars(), "dR"));
            putBits(f, recoVar(tc->getRecoVars(), "innerR"));
            std::fprintf(f, "\n");
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s dump|time workload geomdir event0 nevents out|threads [stages]\n", argv[0]);
        return 2;
    }
    std::string mode = argv[1];
    int workload = std::atoi(argv[2]);
    std::string geomdir = argv[3];
    long e0 = std::atol(argv[4]);
    long ne = std::atol(argv[5]);
    sdlgen::Geometry g;
    loadGeometry(workload, geomdir, g);
    sdlgen::EventConfig cfg = sdlgen::presetConfig(workload);

    if (mode == "dumppseudo") {  // pseudo mini-doublets, then the module-map chain
        FILE* f = std::fopen(argv[6], "w");
        if (!f) return 3;
        for (long ie = e0; ie < e0 + ne; ++ie) {
            sdlgen::Event ev;
            sdlgen::generateEvent(g, cfg, ie, ev);
            SDL::Event sev;
            fill(sev, ev);
            sev.createPseudoMiniDoubletsFromAnchorModule();
            sev.createSegmentsWithModuleMap();
            sev.createTriplets();
            sev.createTrackletsWithModuleMap();
            sev.createTrackCandidatesFromTriplets();
            sev.createTrackCandidatesFromTracklets();
            dumpEvent(f, sev, ie);
            dumpLayerCounters(f, sev);
        }
        std::fclose(f);
        return 0;
    }
    if (mode == "dumpnm1") {  // the module-map chain with createTrackletsWithModuleMap(DefaultNm1_TLAlgo)
        FILE* f = std::fopen(argv[6], "w");
        if (!f) return 3;
        for (long ie = e0; ie < e0 + ne; ++ie) {
            sdlgen::Event ev;
            sdlgen::generateEvent(g, cfg, ie, ev);
            SDL::Event sev;
            fill(sev, ev);
            sev.createMiniDoublets();
            sev.createSegmentsWithModuleMap();
            sev.createTriplets();
            sev.createTrackletsWithModuleMap(SDL::DefaultNm1_TLAlgo);
            sev.createTrackCandidatesFromTriplets();
            sev.createTrackCandidatesFromTracklets();
            dumpEvent(f, sev, ie);
        }
        std::fclose(f);
        return 0;
    }
    if (mode == "dumpnav") {  // createMiniDoublets, createSegmentsWithModuleMap, createTrackletsViaNavigation
        FILE* f = std::fopen(argv[6], "w");
        if (!f) return 3;
        for (long ie = e0; ie < e0 + ne; ++ie) {
            sdlgen::Event ev;
            sdlgen::generateEvent(g, cfg, ie, ev);
            SDL::Event sev;
            fill(sev, ev);
            sev.createMiniDoublets();
            sev.createSegmentsWithModuleMap();
            sev.createTrackletsViaNavigation();
            dumpLayerEvent(f, sev, ie);
        }
        std::fclose(f);
        return 0;
    }
    if (mode == "dumplayers") {
        FILE* f = std::fopen(argv[6], "w");
        if (!f) return 3;
        for (long ie = e0; ie < e0 + ne; ++ie) {
            sdlgen::Event ev;
            sdlgen::generateEvent(g, cfg, ie, ev);
            SDL::Event sev;
            fill(sev, ev);
            runLayerChain(sev);
            dumpLayerEvent(f, sev, ie);
        }
        std::fclose(f);
        return 0;
    }
    if (mode == "dumpvars") {
        FILE* f = std::fopen(argv[6], "w");
        if (!f) return 3;
        for (long ie = e0; ie < e0 + ne; ++ie) {
            sdlgen::Event ev;
            sdlgen::generateEvent(g, cfg, ie, ev);
            SDL::Event sev;
            fill(sev, ev);
            runChain(sev, 6);
            dumpVars(f, sev, ie);
        }
        std::fclose(f);
        return 0;
    }
    if (mode == "dumpseq") {  // any sequence of create* calls with their algorithms
        if (argc < 8) return 2;
        std::vector<std::pair<std::string, int>> seq;
        std::string spec = argv[7];
        size_t pos = 0;
        while (pos < spec.size()) {
            size_t comma = spec.find(',', pos);
            if (comma == std::string::npos) comma = spec.size();
            const std::string tok = spec.substr(pos, comma - pos);
            const size_t colon = tok.find(':');
            if (colon == std::string::npos) return 2;
            seq.emplace_back(tok.substr(0, colon), std::atoi(tok.c_str() + colon + 1));
            pos = comma + 1;
        }
        FILE* f = std::fopen(argv[6], "w");
        if (!f) return 3;
        for (long ie = e0; ie < e0 + ne; ++ie) {
            sdlgen::Event ev;
            sdlgen::generateEvent(g, cfg, ie, ev);
            SDL::Event sev;
            fill(sev, ev);
            for (auto& t : seq)
                if (!runToken(sev, t.first, t.second)) return 2;
            dumpSeqEvent(f, sev, ie);
        }
        std::fclose(f);
        return 0;
    }
    if (mode == "dump") {
        FILE* f = std::fopen(argv[6], "w");
        if (!f) return 3;
        for (long ie = e0; ie < e0 + ne; ++ie) {
            sdlgen::Event ev;
            sdlgen::generateEvent(g, cfg, ie, ev);
            SDL::Event sev;
            fill(sev, ev);
            runChain(sev, 6);
            dumpEvent(f, sev, ie);
        }
        std::fclose(f);
        return 0;
    }
    if (mode == "time") {
        int threads = std::atoi(argv[6]);
        int stages = argc > 7 ? std::atoi(argv[7]) : 6;
        std::vector<sdlgen::Event> evs(ne);
        for (long i = 0; i < ne; ++i) sdlgen::generateEvent(g, cfg, e0 + i, evs[i]);
        std::atomic<long> next(0);
        std::vector<double> perEvent(ne, 0.0);
        auto t0 = std::chrono::steady_clock::now();
        auto work = [&]() {
            for (;;) {
                long i = next.fetch_add(1);
                if (i >= ne) break;
                auto a = std::chrono::steady_clock::now();
                SDL::Event sev;
                fill(sev, evs[i]);
                runChain(sev, stages);
                auto b = std::chrono::steady_clock::now();
                perEvent[i] = std::chrono::duration<double>(b - a).count();
            }
        };
        std::vector<std::thread> th;
        for (int t = 0; t < threads; ++t) th.emplace_back(work);
        for (auto& t : th) t.join();
        double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        double sum = 0;
        for (double x : perEvent) sum += x;
        std::printf("{\"events\": %ld, \"threads\": %d, \"wall_s\": %.6f, \"events_per_s\": %.6f, \"mean_event_s\": %.6f}\n",
                    ne, threads, wall, ne / wall, sum / ne);
        return 0;
    }
    return 2;
}
