// SDL::EndcapGeometry of the shim (the reference's EndcapGeometry.h:12-49):
// "detid avgr2 yl sl yh sh c7 c8 cz" rows (EndcapGeometry.cc:18-61), with the
// reference's reading of columns 7 and 8 (column 7 is returned as the centroid
// phi, column 8 as the centroid r).  A detId without a row reads 0.
#ifndef SDL_AMD_ENDCAPGEOMETRY_H
#define SDL_AMD_ENDCAPGEOMETRY_H

#include <array>
#include <map>
#include <string>

namespace SDL {

class EndcapGeometry {
  public:
    EndcapGeometry() = default;
    EndcapGeometry(std::string filename) { load(filename); }
    void load(std::string filename);
    float getAverageR2(unsigned int detid) const { return col(detid, 0); }
    float getYInterceptLower(unsigned int detid) const { return col(detid, 1); }
    float getSlopeLower(unsigned int detid) const { return col(detid, 2); }
    float getYInterceptUpper(unsigned int detid) const { return col(detid, 3); }
    float getSlopeUpper(unsigned int detid) const { return col(detid, 4); }
    float getCentroidR(unsigned int detid) const { return col(detid, 6); }
    float getCentroidPhi(unsigned int detid) const { return col(detid, 5); }
    float getCentroidZ(unsigned int detid) const { return col(detid, 7); }
    const std::string& file() const { return file_; }
    // the eight columns after the detId, as read from the file
    const std::map<unsigned int, std::array<float, 8>>& rows() const { return rows_; }
    unsigned version() const { return version_; }

  private:
    float col(unsigned int detid, int k) const;
    std::string file_;
    std::map<unsigned int, std::array<float, 8>> rows_;
    unsigned version_ = 0;
};
extern EndcapGeometry endcapGeometry;

}  // namespace SDL

#endif
