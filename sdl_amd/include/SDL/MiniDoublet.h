// SDL::MiniDoublet of the shim (the reference's MiniDoublet.h:17-153): a stored
// mini-doublet as the device found it, with the values its algorithm set
// (the typed getters of MiniDoublet.h:79-85 and the reco-variable map).
#ifndef SDL_AMD_MINIDOUBLET_H
#define SDL_AMD_MINIDOUBLET_H

#include <map>
#include <string>
#include <vector>

namespace SDL {

class Hit;
class Segment;

class MiniDoublet {
  public:
    MiniDoublet(Hit* lowerHit, Hit* upperHit, Hit* anchorHit);
    Hit* lowerHitPtr() const { return lowerHitPtr_; }
    Hit* upperHitPtr() const { return upperHitPtr_; }
    Hit* anchorHitPtr() const { return anchorHitPtr_; }
    const int& getPassAlgo() const { return passAlgo_; }
    const float& getDz() const { return v_[0]; }
    const float& getShiftedDz() const { return v_[1]; }
    const float& getDeltaPhi() const { return v_[2]; }
    const float& getDeltaPhiChange() const { return v_[3]; }
    const float& getDeltaPhiNoShift() const { return v_[4]; }
    const float& getDeltaPhiChangeNoShift() const { return v_[5]; }
    const std::map<std::string, float>& getRecoVars() const { return recovars_; }
    const float& getRecoVar(std::string key) const { return recovars_.at(key); }
    // MiniDoublet::isIdxMatched (MiniDoublet.cc:686-693)
    bool isIdxMatched(const MiniDoublet&) const;
    // the segments of the stored segment list that start / end here
    const std::vector<Segment*>& getListOfOutwardSegmentPtrs() { return outward_; }
    const std::vector<Segment*>& getListOfInwardSegmentPtrs() { return inward_; }

    // set by the event when it builds its views
    void setValues(const float* v, int passAlgo);
    void setRecoVars(const std::string& key, float value) { recovars_[key] = value; }
    void addOutwardSegmentPtr(Segment* sg) { outward_.push_back(sg); }
    void addInwardSegmentPtr(Segment* sg) { inward_.push_back(sg); }

  private:
    Hit *lowerHitPtr_, *upperHitPtr_, *anchorHitPtr_;
    int passAlgo_ = 0;
    float v_[6] = {0, 0, 0, 0, 0, 0};
    std::map<std::string, float> recovars_;
    std::vector<Segment*> outward_, inward_;
};

}  // namespace SDL

#endif
