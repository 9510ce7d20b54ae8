// SDL::Tracklet of the shim (the reference's Tracklet.h:15-139): two segments
// four layers deep, with the typed values of Tracklet.h:85-90.
#ifndef SDL_AMD_TRACKLET_H
#define SDL_AMD_TRACKLET_H

#include "TrackletBase.h"

namespace SDL {

class Tracklet : public TrackletBase {
  public:
    Tracklet(Segment* innerSegmentPtr, Segment* outerSegmentPtr) : TrackletBase(innerSegmentPtr, outerSegmentPtr) {}
    const float& getBetaIn() const { return v_[0]; }
    const float& getBetaInCut() const { return v_[1]; }
    const float& getBetaOut() const { return v_[2]; }
    const float& getBetaOutCut() const { return v_[3]; }
    const float& getDeltaBeta() const { return v_[4]; }
    const float& getDeltaBetaCut() const { return v_[5]; }
    // Tracklet::hasCommonSegment (Tracklet.cc:3326-3331)
    bool hasCommonSegment(const Tracklet&) const;

    void setValues(const float* v) {
        for (int k = 0; k < 6; ++k) v_[k] = v[k];
    }

  private:
    float v_[6] = {};
};

}  // namespace SDL

#endif
