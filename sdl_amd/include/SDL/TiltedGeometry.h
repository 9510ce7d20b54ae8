// SDL::TiltedGeometry of the shim (the reference's TiltedGeometry.h:12-37):
// "detid drdz slope" rows (TiltedGeometry.cc:18-41).  A detId without a row
// reads 0, as the reference's std::map lookups return.
#ifndef SDL_AMD_TILTEDGEOMETRY_H
#define SDL_AMD_TILTEDGEOMETRY_H

#include <map>
#include <string>

namespace SDL {

class TiltedGeometry {
  public:
    TiltedGeometry() = default;
    TiltedGeometry(std::string filename) { load(filename); }
    void load(std::string filename);
    float getDrDz(unsigned int detid) const;
    float getSlope(unsigned int detid) const;
    const std::string& file() const { return file_; }
    const std::map<unsigned int, float>& drdzs() const { return drdzs_; }
    const std::map<unsigned int, float>& slopes() const { return slopes_; }
    unsigned version() const { return version_; }

  private:
    std::string file_;
    std::map<unsigned int, float> drdzs_, slopes_;
    unsigned version_ = 0;
};
extern TiltedGeometry tiltedGeometry;

}  // namespace SDL

#endif
