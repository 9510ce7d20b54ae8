// SDL::Event over the MI355X C-ABI (include/sdl_amd.h): a drop-in for the
// reference's host API (Event.h:24-411 with Hit.h, Module.h, Layer.h,
// MiniDoublet.h, Segment.h, Triplet.h, Tracklet.h, TrackCandidate.h,
// TiltedGeometry.h, EndcapGeometry.h, ModuleConnectionMap.h and Algo.h).
//
// Differences a caller can see:
//  * the device builds the geometry from the three singletons when the first
//    Event is created after a load() or add();
//  * the objects (MiniDoublet, Segment, Triplet, Tracklet, TrackCandidate) are
//    host views of the device results, built when they are first read after a
//    stage ran (getModulePtrs(), getLayer(), ...): the same objects, with the
//    same pointers between them and the same values as the reference's, but
//    inside a module or layer list not necessarily in the reference's order;
//    the strip-shifted hits of a mini-doublet (getLowerShiftedHit) are not kept;
//  * an algorithm value that Algo.h does not define is rejected with
//    std::invalid_argument (the reference warns and stores nothing);
//    createTrackCandidatesFromTriplets after AllComb segments or triplets
//    throws std::logic_error;
//  * the objects can also be read as hit-index tuples (getMiniDoublets() etc.)
//    and their values as float rows parallel to them (getMiniDoubletVariables() etc.).
#ifndef SDL_AMD_EVENT_H
#define SDL_AMD_EVENT_H

#include <array>
#include <deque>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "Algo.h"
#include "EndcapGeometry.h"
#include "Hit.h"
#include "Layer.h"
#include "MiniDoublet.h"
#include "Module.h"
#include "ModuleConnectionMap.h"
#include "Segment.h"
#include "TiltedGeometry.h"
#include "TrackCandidate.h"
#include "Tracklet.h"
#include "Triplet.h"

struct sdl_event;

namespace SDL {

class Event {
  public:
    Event();
    ~Event();
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;

    void setLogLevel(LogLevel logLevel = Log_Nothing) { logLevel_ = logLevel; }

    // Event.h:197-204: modules are created as hits arrive (and a lower module's
    // partner by the mini-doublet stages, Event.cc:262); getModule creates a
    // module that does not exist yet, as the reference's does (Event.cc:54-78)
    bool hasModule(unsigned int detId);
    Module& getModule(unsigned int detId);
    const std::vector<Module*> getModulePtrs();
    const std::vector<Module*> getLowerModulePtrs();
    // barrel layers 1-6 and endcap layers 1-5 (Event.cc:90-118)
    Layer& getLayer(int layer, Layer::SubDet subdet);
    const std::vector<Layer*> getLayerPtrs();

    // Event.cc:120-139 (the hit is buffered on the host until the first stage runs)
    void addHitToModule(Hit hit, unsigned int detId);

    // Event.h:240-300, module-map family
    void createMiniDoublets(MDAlgo algo = Default_MDAlgo);
    // Event.cc:307-430: one pseudo mini-doublet per anchor-side hit (with the
    // partner sensor's first hit), instead of createMiniDoublets.
    void createPseudoMiniDoubletsFromAnchorModule(MDAlgo algo = Default_MDAlgo);
    void createSegmentsWithModuleMap(SGAlgo algo = Default_SGAlgo);
    void createTriplets(TPAlgo algo = Default_TPAlgo);
    void createTrackletsWithModuleMap(TLAlgo algo = Default_TLAlgo);
    void createTrackCandidatesFromTriplets(TCAlgo algo = Default_TCAlgo);
    void createTrackCandidatesFromTracklets(TCAlgo algo = Default_TCAlgo);

    // Event.h:240-300, layer-pair family (Event.cc:432-443, 599-609, 823-837):
    // every object of an inner layer against every object of an outer layer
    // for the reference's fixed layer pairs; run after createMiniDoublets,
    // after either segment stage and after either tracklet stage respectively.
    // Their objects are returned by getSegments, getTracklets and
    // getTrackCandidatesFromTracklets.
    void createSegments(SGAlgo algo = Default_SGAlgo);
    void createTracklets(TLAlgo algo = Default_TLAlgo);
    void createTrackCandidates(TCAlgo algo = Default_TCAlgo);
    // Event.cc:696-751: tracklets (inner SG, SG two segments further out along the
    // MDs' outward segment lists); returned by getTracklets.
    void createTrackletsViaNavigation(TLAlgo algo = Default_TLAlgo);

    // Event.h:302-405
    unsigned int getNumberOfHits();
    unsigned int getNumberOfHitsByLayerBarrel(unsigned int);
    unsigned int getNumberOfHitsByLayerEndcap(unsigned int);
    unsigned int getNumberOfHitsByLayerBarrelUpperModule(unsigned int);
    unsigned int getNumberOfHitsByLayerEndcapUpperModule(unsigned int);
    unsigned int getNumberOfMiniDoublets();
    unsigned int getNumberOfSegments();
    unsigned int getNumberOfTracklets();
    unsigned int getNumberOfTriplets();
    unsigned int getNumberOfTrackCandidates();
    unsigned int getNumberOfMiniDoubletCandidates();
    unsigned int getNumberOfSegmentCandidates();
    unsigned int getNumberOfTrackletCandidates();
    unsigned int getNumberOfTripletCandidates();
    unsigned int getNumberOfTrackCandidateCandidates();
    unsigned int getNumberOfMiniDoubletCandidatesByLayerBarrel(unsigned int);
    unsigned int getNumberOfSegmentCandidatesByLayerBarrel(unsigned int);
    unsigned int getNumberOfTrackletCandidatesByLayerBarrel(unsigned int);
    unsigned int getNumberOfTripletCandidatesByLayerBarrel(unsigned int);
    unsigned int getNumberOfTrackCandidateCandidatesByLayerBarrel(unsigned int);
    unsigned int getNumberOfMiniDoubletCandidatesByLayerEndcap(unsigned int);
    unsigned int getNumberOfSegmentCandidatesByLayerEndcap(unsigned int);
    unsigned int getNumberOfTrackletCandidatesByLayerEndcap(unsigned int);
    unsigned int getNumberOfTripletCandidatesByLayerEndcap(unsigned int);
    unsigned int getNumberOfTrackCandidateCandidatesByLayerEndcap(unsigned int);
    unsigned int getNumberOfMiniDoubletsByLayerBarrel(unsigned int);
    unsigned int getNumberOfSegmentsByLayerBarrel(unsigned int);
    unsigned int getNumberOfTrackletsByLayerBarrel(unsigned int);
    unsigned int getNumberOfTripletsByLayerBarrel(unsigned int);
    unsigned int getNumberOfTrackCandidatesByLayerBarrel(unsigned int);
    unsigned int getNumberOfMiniDoubletsByLayerEndcap(unsigned int);
    unsigned int getNumberOfSegmentsByLayerEndcap(unsigned int);
    unsigned int getNumberOfTrackletsByLayerEndcap(unsigned int);
    unsigned int getNumberOfTripletsByLayerEndcap(unsigned int);
    unsigned int getNumberOfTrackCandidatesByLayerEndcap(unsigned int);

    // Objects as hit-index tuples: MD (lower, upper), SG = two MDs, TP / TL = two
    // SGs, TC = two TPs (from triplets) or two TLs (from tracklets).
    std::vector<std::array<int, 2>> getMiniDoublets();
    std::vector<std::array<int, 4>> getSegments();
    std::vector<std::array<int, 8>> getTriplets();
    std::vector<std::array<int, 8>> getTracklets();
    std::vector<std::array<int, 16>> getTrackCandidatesFromTriplets();
    std::vector<std::array<int, 16>> getTrackCandidatesFromTracklets();

    // Per-object values, one row per object in the order of the matching object
    // getter above: the typed getters and reco variables of the reference's
    // MiniDoublet (MiniDoublet.h:85-97, setRecoVars in MiniDoublet.cc:113-197),
    // Segment (Segment.h:128-149, Segment.cc:104-284) and Tracklet
    // (TrackletBase.cc:50-60, Tracklet.cc:43-98).  Column names come from
    // variableNames(); a value a path never sets keeps the reference's default
    // (0 for typed fields, -999 for reco variables, NaN where the reco map has
    // no entry).
    std::vector<std::array<float, 7>> getMiniDoubletVariables();
    std::vector<std::array<float, 22>> getSegmentVariables();
    std::vector<std::array<float, 114>> getTrackletVariables();
    // dR and innerR of each TC of getTrackCandidatesFromTracklets() (TrackCandidate.cc:206-207)
    std::vector<std::array<float, 2>> getTrackCandidatesFromTrackletsVariables();
    // Column names of the rows above for "MD", "SG", "TL" or "TC" (empty otherwise).
    static std::vector<std::string> variableNames(const std::string& objectType);

  private:
    bool run(int stage, int algo, const char* what);
    unsigned int counter(int which, int bin);
    template <int W>
    std::vector<std::array<int, W>> objects(int stage);
    template <int W>
    std::vector<std::array<float, W>> variables(int stage);
    bool exported(int stage) const;
    void buildViews();
    void addPartnerModules();
    Layer& layerOf(const Module& m);

    sdl_event* ev_ = nullptr;
    LogLevel logLevel_ = Log_Nothing;
    unsigned ran_ = 0;  // stages run since the last hit was added (bit = SDL_STAGE_*)
    int algo_[11] = {};  // algorithm each stage last ran with
    bool tlFromLayers_ = false;  // the stored tracklets came from createTracklets (layer pairs)
    unsigned int nHits_ = 0;
    std::array<unsigned int, 11> hitsByLayer_{};       // lower sensors
    std::array<unsigned int, 11> hitsByLayerUpper_{};  // upper sensors
    // the reference's storage: hits, modules (by detId, and in creation order), layers
    std::deque<Hit> hits_, hitEdges_;
    std::unordered_map<int, Hit*> hitByIdx_;
    std::map<unsigned int, Module> modules_;
    std::vector<Module*> modulePtrs_, lowerModulePtrs_;
    std::map<int, Layer> barrelLayers_, endcapLayers_;
    std::vector<Layer*> layerPtrs_;
    // host views of the device's objects (valid while viewsBuilt_)
    bool viewsBuilt_ = false;
    std::deque<MiniDoublet> miniDoublets_;
    std::deque<Segment> segments_;
    std::deque<Triplet> triplets_;
    std::deque<Tracklet> tracklets_;
    std::deque<TrackCandidate> trackCandidates_;
};

}  // namespace SDL

#endif
