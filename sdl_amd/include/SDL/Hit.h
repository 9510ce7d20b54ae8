// SDL::Hit of the shim (the reference's Hit.h:19-93): a point with its index
// and the derived quantities of Hit::setDerivedQuantities (Hit.cc:68-83).
#ifndef SDL_AMD_HIT_H
#define SDL_AMD_HIT_H

#include <iosfwd>

namespace SDL {

class Module;

class Hit {
  public:
    Hit();
    Hit(float x, float y, float z, int idx = -1);
    Hit(const Hit&);
    ~Hit();

    void setX(float x);
    void setY(float y);
    void setZ(float z);
    void setXYZ(float x, float y, float z);
    void setIdx(int idx);
    void setModule(const Module*);
    void setHitHighEdgePtr(Hit* hit);
    void setHitLowEdgePtr(Hit* hit);

    const float& x() const { return x_; }
    const float& y() const { return y_; }
    const float& z() const { return z_; }
    const float& rt() const { return rt_; }
    const float& r3() const { return r3_; }
    const float& phi() const { return phi_; }
    const float& eta() const { return eta_; }
    const int& idx() const { return idx_; }
    // the module the hit was added to (Event::addHitToModule); throws for a hit
    // that was not added to an event
    const Module& getModule() const;
    // the strip edges of an endcap-2S hit (GeometryUtil.cc:3-23), else null
    const Hit* getHitHighEdgePtr() const { return hitHighEdge_; }
    const Hit* getHitLowEdgePtr() const { return hitLowEdge_; }

    // Hit.cc:149-182
    float deltaPhi(const Hit&) const;
    float deltaPhiChange(const Hit&) const;
    bool isIdxMatched(const Hit&) const;

    bool operator!=(const Hit&) const;
    bool operator==(const Hit&) const;
    Hit operator-(const Hit&) const;
    Hit& operator=(const Hit&);
    Hit& operator-=(const Hit&);
    Hit& operator+=(const Hit&);
    Hit& operator/=(const float&);

    friend std::ostream& operator<<(std::ostream& out, const Hit& hit);
    friend std::ostream& operator<<(std::ostream& out, const Hit* hit);

  private:
    void setDerivedQuantities();
    float x_, y_, z_, rt_, r3_, phi_, eta_;
    int idx_;
    const Module* module_ = nullptr;
    Hit* hitHighEdge_ = nullptr;
    Hit* hitLowEdge_ = nullptr;
};

}  // namespace SDL

#endif
