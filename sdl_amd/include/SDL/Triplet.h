// SDL::Triplet of the shim (the reference's Triplet.h:14-81): two segments that
// share a mini-doublet.
#ifndef SDL_AMD_TRIPLET_H
#define SDL_AMD_TRIPLET_H

#include "TrackletBase.h"

namespace SDL {

class Triplet : public TrackletBase {
  public:
    Triplet(Segment* innerSegmentPtr, Segment* outerSegmentPtr) : TrackletBase(innerSegmentPtr, outerSegmentPtr) {}
};

}  // namespace SDL

#endif
