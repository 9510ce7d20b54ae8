// SDL::Layer of the shim (the reference's Layer.h:17-124): the objects stored
// at one barrel or endcap layer (host views of the device results), and the
// reference's compatible layer pairs (Layer.cc:3-52).
#ifndef SDL_AMD_LAYER_H
#define SDL_AMD_LAYER_H

#include <iosfwd>
#include <utility>
#include <vector>

namespace SDL {

class MiniDoublet;
class Segment;
class Triplet;
class Tracklet;
class TrackCandidate;

class Layer {
  public:
    enum BarrelLayers { BarrelLayer0 = 1, BarrelLayer1, BarrelLayer2, BarrelLayer3, BarrelLayer4, BarrelLayer5, nBarrelLayer };
    enum EndcapLayers { EndcapLayer0 = 1, EndcapLayer1, EndcapLayer2, EndcapLayer3, EndcapLayer4, EndcapLayer5, nEndcapLayer };
    enum SubDet { Barrel = 5, Endcap = 4 };
    typedef std::vector<std::pair<std::pair<int, SubDet>, std::pair<int, SubDet>>> LayerPairs;

    Layer();
    Layer(int layerIdx, unsigned short subdet);
    ~Layer();

    const unsigned short& subdet() const { return subdet_; }
    const int& layerIdx() const { return layerIdx_; }
    const std::vector<MiniDoublet*>& getMiniDoubletPtrs() const { return miniDoublets_; }
    const std::vector<Segment*>& getSegmentPtrs() const { return segments_; }
    const std::vector<Triplet*>& getTripletPtrs() const { return triplets_; }
    const std::vector<Tracklet*>& getTrackletPtrs() const { return tracklets_; }
    const std::vector<TrackCandidate*>& getTrackCandidatePtrs() const { return trackCandidates_; }
    static const LayerPairs& getListOfTrackletCompatibleLayerPairs();
    static const LayerPairs& getListOfSegmentCompatibleLayerPairs();

    void setLayerIdx(int lidx) { layerIdx_ = lidx; }
    void setSubDet(SubDet subdet) { subdet_ = subdet; }
    void addMiniDoublet(MiniDoublet* md) { miniDoublets_.push_back(md); }
    void addSegment(Segment* sg) { segments_.push_back(sg); }
    void addTriplet(Triplet* tp) { triplets_.push_back(tp); }
    void addTracklet(Tracklet* tl) { tracklets_.push_back(tl); }
    void addTrackCandidate(TrackCandidate* tc) { trackCandidates_.push_back(tc); }
    void clearObjects();

    friend std::ostream& operator<<(std::ostream& os, const Layer& layer);
    friend std::ostream& operator<<(std::ostream& os, const Layer* layer);

  private:
    int layerIdx_ = 0;
    unsigned short subdet_ = 0;
    std::vector<MiniDoublet*> miniDoublets_;
    std::vector<Segment*> segments_;
    std::vector<Triplet*> triplets_;
    std::vector<Tracklet*> tracklets_;
    std::vector<TrackCandidate*> trackCandidates_;
};

}  // namespace SDL

#endif
