// SDL::Segment of the shim (the reference's Segment.h:17-188): a stored segment
// with its two mini-doublets and the values its algorithm set
// (Segment.h:110-128 getters and the reco-variable map).
#ifndef SDL_AMD_SEGMENT_H
#define SDL_AMD_SEGMENT_H

#include <map>
#include <string>
#include <vector>

namespace SDL {

class MiniDoublet;
class Tracklet;

class Segment {
  public:
    Segment(MiniDoublet* innerMiniDoubletPtr, MiniDoublet* outerMiniDoubletPtr);
    MiniDoublet* innerMiniDoubletPtr() const { return inner_; }
    MiniDoublet* outerMiniDoubletPtr() const { return outer_; }
    const int& getPassAlgo() const { return passAlgo_; }
    const float& getRtIn() const { return v_[0]; }
    const float& getRtOut() const { return v_[1]; }
    const float& getZIn() const { return v_[2]; }
    const float& getZOut() const { return v_[3]; }
    const float& getZLo() const { return v_[4]; }
    const float& getZHi() const { return v_[5]; }
    const float& getRtLo() const { return v_[6]; }
    const float& getRtHi() const { return v_[7]; }
    const float& getDeltaPhi() const { return v_[8]; }
    const float& getDeltaPhiMin() const { return v_[9]; }
    const float& getDeltaPhiMax() const { return v_[10]; }
    const float& getDeltaPhiChange() const { return v_[11]; }
    const float& getDeltaPhiMinChange() const { return v_[12]; }
    const float& getDeltaPhiMaxChange() const { return v_[13]; }
    const std::map<std::string, float>& getRecoVars() const { return recovars_; }
    const float& getRecoVar(std::string key) const { return recovars_.at(key); }
    // Segment::hasCommonMiniDoublet, isIdxMatched (Segment.cc:861-875)
    bool hasCommonMiniDoublet(const Segment&) const;
    bool isIdxMatched(const Segment&) const;
    const std::vector<Tracklet*>& getListOfOutwardTrackletPtrs() { return outward_; }
    const std::vector<Tracklet*>& getListOfInwardTrackletPtrs() { return inward_; }

    void setValues(const float* v, int passAlgo);
    void setRecoVars(const std::string& key, float value) { recovars_[key] = value; }
    void addOutwardTrackletPtr(Tracklet* tl) { outward_.push_back(tl); }
    void addInwardTrackletPtr(Tracklet* tl) { inward_.push_back(tl); }

  private:
    MiniDoublet *inner_, *outer_;
    int passAlgo_ = 0;
    float v_[14] = {};
    std::map<std::string, float> recovars_;
    std::vector<Tracklet*> outward_, inward_;
};

}  // namespace SDL

#endif
