// SDL::TrackletBase of the shim (the reference's TrackletBase.h:14-69): a pair
// of segments, the base of Triplet and Tracklet.
#ifndef SDL_AMD_TRACKLETBASE_H
#define SDL_AMD_TRACKLETBASE_H

#include <map>
#include <string>

namespace SDL {

class Segment;

class TrackletBase {
  public:
    TrackletBase(Segment* innerSegmentPtr, Segment* outerSegmentPtr);
    virtual ~TrackletBase();
    Segment* innerSegmentPtr() const { return inner_; }
    Segment* outerSegmentPtr() const { return outer_; }
    const int& getPassAlgo() const { return passAlgo_; }
    const std::map<std::string, float>& getRecoVars() const { return recovars_; }
    const float& getRecoVar(std::string key) const { return recovars_.at(key); }
    // TrackletBase::isIdxMatched (TrackletBase.cc:63-70)
    bool isIdxMatched(const TrackletBase&) const;

    void setPassAlgo(int passAlgo) { passAlgo_ = passAlgo; }
    void setRecoVars(const std::string& key, float value) { recovars_[key] = value; }

  protected:
    Segment *inner_, *outer_;
    int passAlgo_ = 0;
    std::map<std::string, float> recovars_;
};

}  // namespace SDL

#endif
