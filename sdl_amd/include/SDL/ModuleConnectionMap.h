// SDL::ModuleConnectionMap of the shim (the reference's ModuleConnectionMap.h:12-36):
// "detid n d1 ... dn" rows.  load() replaces the map with a file's lists in
// file order (ModuleConnectionMap.cc:18-47); add() merges a file's lists into
// it, sorted with duplicates removed (ModuleConnectionMap.cc:49-85).  A detId
// without a row has an empty list.
#ifndef SDL_AMD_MODULECONNECTIONMAP_H
#define SDL_AMD_MODULECONNECTIONMAP_H

#include <map>
#include <string>
#include <vector>

namespace SDL {

class ModuleConnectionMap {
  public:
    ModuleConnectionMap() = default;
    ModuleConnectionMap(std::string filename) { load(filename); }
    void load(std::string filename);
    void add(std::string filename);
    void print() const;
    const std::vector<unsigned int>& getConnectedModuleDetIds(unsigned int detid) const;
    const std::string& file() const { return file_; }
    const std::map<unsigned int, std::vector<unsigned int>>& connections() const { return conn_; }
    unsigned version() const { return version_; }

  private:
    std::string file_;
    std::map<unsigned int, std::vector<unsigned int>> conn_;
    unsigned version_ = 0;
};
extern ModuleConnectionMap moduleConnectionMap;

}  // namespace SDL

#endif
