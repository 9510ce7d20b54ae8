// SDL::Module of the shim (the reference's Module.h:40-300): the decoded detId
// (through the library's sdl_module_decode, Module.cc:159-358) and the lists of
// the hits and objects stored at the module.  The object lists are host views
// of the device results, filled when the event's objects are first read.
#ifndef SDL_AMD_MODULE_H
#define SDL_AMD_MODULE_H

#include <iosfwd>
#include <vector>

namespace SDL {

class Hit;
class MiniDoublet;
class Segment;
class Triplet;
class Tracklet;

class Module {
  public:
    enum SubDet { Barrel = 5, Endcap = 4 };
    enum Side { NegZ = 1, PosZ = 2, Center = 3 };
    enum ModuleType { PS, TwoS };
    enum ModuleLayerType { Pixel, Strip };

    Module();
    Module(unsigned int detId);
    Module(const Module&);
    ~Module();

    const unsigned int& detId() const { return detId_; }
    const unsigned int& partnerDetId() const { return partnerDetId_; }
    const unsigned short& subdet() const { return subdet_; }
    const unsigned short& side() const { return side_; }
    const unsigned short& layer() const { return layer_; }
    const unsigned short& rod() const { return rod_; }
    const unsigned short& ring() const { return ring_; }
    const unsigned short& module() const { return module_; }
    const unsigned short& isLower() const { return isLower_; }
    const bool& isInverted() const { return isInverted_; }
    const ModuleType& moduleType() const { return moduleType_; }
    const ModuleLayerType& moduleLayerType() const { return moduleLayerType_; }
    const std::vector<Hit*>& getHitPtrs() const { return hits_; }
    const std::vector<MiniDoublet*>& getMiniDoubletPtrs() const { return miniDoublets_; }
    const std::vector<Segment*>& getSegmentPtrs() const { return segments_; }
    const std::vector<Triplet*>& getTripletPtrs() const { return triplets_; }
    const std::vector<Tracklet*>& getTrackletPtrs() const { return tracklets_; }

    void setDetId(unsigned int);
    void addHit(Hit* hit);
    void addMiniDoublet(MiniDoublet* md);
    void addSegment(Segment* sg);
    void addTriplet(Triplet* tp);
    void addTracklet(Tracklet* tl);
    void clearObjects();  // the object lists (not the hits), before the views are rebuilt

    static unsigned short parseSubdet(unsigned int);
    static unsigned short parseSide(unsigned int);
    static unsigned short parseLayer(unsigned int);
    static unsigned short parseRod(unsigned int);
    static unsigned short parseRing(unsigned int);
    static unsigned short parseModule(unsigned int);
    static unsigned short parseIsLower(unsigned int);
    static bool parseIsInverted(unsigned int);
    static unsigned int parsePartnerDetId(unsigned int);
    static ModuleType parseModuleType(unsigned int);
    static ModuleLayerType parseModuleLayerType(unsigned int);

    friend std::ostream& operator<<(std::ostream& os, const Module& module);
    friend std::ostream& operator<<(std::ostream& os, const Module* module);

  private:
    unsigned int detId_ = 0, partnerDetId_ = 0;
    unsigned short subdet_ = 0, side_ = 0, layer_ = 0, rod_ = 0, ring_ = 0, module_ = 0, isLower_ = 0;
    bool isInverted_ = false;
    ModuleType moduleType_ = PS;
    ModuleLayerType moduleLayerType_ = Pixel;
    std::vector<Hit*> hits_;
    std::vector<MiniDoublet*> miniDoublets_;
    std::vector<Segment*> segments_;
    std::vector<Triplet*> triplets_;
    std::vector<Tracklet*> tracklets_;
};

}  // namespace SDL

#endif
