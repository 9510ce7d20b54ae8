// SDL::TrackCandidate of the shim (the reference's TrackCandidate.h:17-127):
// two triplets (TCs from triplets) or two tracklets (TCs from tracklets).  The
// typed accessors throw std::logic_error for the other kind, as the
// reference's do (TrackCandidate.cc:38-88).
#ifndef SDL_AMD_TRACKCANDIDATE_H
#define SDL_AMD_TRACKCANDIDATE_H

#include <map>
#include <string>

namespace SDL {

class TrackletBase;
class Tracklet;
class Triplet;

class TrackCandidate {
  public:
    TrackCandidate(TrackletBase* innerTrackletPtr, TrackletBase* outerTrackletPtr);
    TrackletBase* innerTrackletBasePtr() const { return inner_; }
    TrackletBase* outerTrackletBasePtr() const { return outer_; }
    Tracklet* innerTrackletPtr() const;
    Tracklet* outerTrackletPtr() const;
    Triplet* innerTripletPtr() const;
    Triplet* outerTripletPtr() const;
    const int& getPassAlgo() const { return passAlgo_; }
    const std::map<std::string, float>& getRecoVars() const { return recovars_; }
    const float& getRecoVar(std::string key) const { return recovars_.at(key); }

    void setPassAlgo(int passAlgo) { passAlgo_ = passAlgo; }
    void setRecoVars(const std::string& key, float value) { recovars_[key] = value; }

  private:
    TrackletBase *inner_, *outer_;
    int passAlgo_ = 0;
    std::map<std::string, float> recovars_;
};

}  // namespace SDL

#endif
