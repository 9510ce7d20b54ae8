// SDL::Event over the MI355X C-ABI (include/sdl_amd.h): a drop-in for the
// reference's host API (Event.h:24-411, Hit.h, TiltedGeometry.h,
// EndcapGeometry.h, ModuleConnectionMap.h) for the module-map chain.
//
// Differences a caller can see:
//  * the three geometry singletons only record their file; the device copy of
//    the geometry is built when the first Event is created after a load();
//  * object graphs (Module / Layer / MiniDoublet pointers) are not materialised:
//    objects are read back as hit-index tuples (getMiniDoublets() etc.);
//  * only the Default algorithms, and DefaultNm1_TLAlgo for the three tracklet
//    stages, run on the device; another algorithm value prints a warning and
//    leaves the stage empty, as the reference does for an unrecognised value
//    (MiniDoublet.cc:216-220);
//  * per-object values (the objects' getters and reco variables) are read back
//    as float rows parallel to the object tuples (getMiniDoubletVariables() etc.).
#ifndef SDL_AMD_EVENT_H
#define SDL_AMD_EVENT_H

#include <array>
#include <cmath>
#include <string>
#include <vector>

#include "Algo.h"

struct sdl_event;

namespace SDL {

// Hit.h:19-93: a point with an index; derived quantities on demand.
class Hit {
  public:
    Hit() : x_(0), y_(0), z_(0), idx_(-1) {}
    Hit(float x, float y, float z, int idx = -1) : x_(x), y_(y), z_(z), idx_(idx) {}
    const float& x() const { return x_; }
    const float& y() const { return y_; }
    const float& z() const { return z_; }
    const int& idx() const { return idx_; }
    void setIdx(int idx) { idx_ = idx; }
    float rt() const { return float(std::sqrt(double(x_ * x_ + y_ * y_))); }
    float r3() const { return float(std::sqrt(double(x_ * x_ + y_ * y_ + z_ * z_))); }

  private:
    float x_, y_, z_;
    int idx_;
};

// The three geometry singletons of the reference, recording their source file.
class TiltedGeometry {
  public:
    void load(std::string filename);
    const std::string& file() const { return file_; }
  private:
    std::string file_;
};
class EndcapGeometry {
  public:
    void load(std::string filename);
    const std::string& file() const { return file_; }
  private:
    std::string file_;
};
class ModuleConnectionMap {
  public:
    void load(std::string filename);
    const std::string& file() const { return file_; }
  private:
    std::string file_;
};
extern TiltedGeometry tiltedGeometry;
extern EndcapGeometry endcapGeometry;
extern ModuleConnectionMap moduleConnectionMap;

class Event {
  public:
    Event();
    ~Event();
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;

    void setLogLevel(LogLevel logLevel = Log_Nothing) { logLevel_ = logLevel; }

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

    sdl_event* ev_ = nullptr;
    LogLevel logLevel_ = Log_Nothing;
    unsigned ran_ = 0;  // stages run since the last hit was added (bit = SDL_STAGE_*)
    unsigned int nHits_ = 0;
    std::array<unsigned int, 11> hitsByLayer_{};       // lower sensors
    std::array<unsigned int, 11> hitsByLayerUpper_{};  // upper sensors
};

}  // namespace SDL

#endif
