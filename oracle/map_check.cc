// Test infrastructure: the compiled reference's ModuleConnectionMap
// (oracle/_ref/libsdl_ref.so) loads one map file, merges a second one with
// add() (ModuleConnectionMap.cc:18-85) and writes the result in the files'
// "detid n d1 ... dn" format, one line per detId of either file in increasing
// order, each list as getConnectedModuleDetIds returns it.
//
// Usage: map_check <load-file> <add-file> <out-file>
#include <cstdio>
#include <fstream>
#include <set>
#include <sstream>
#include <string>

#include "ModuleConnectionMap.h"

namespace {
void keysOf(const char* path, std::set<unsigned>& keys) {
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        unsigned d;
        if (ss >> d) keys.insert(d);
    }
}
}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    SDL::ModuleConnectionMap m;
    m.load(argv[1]);
    m.add(argv[2]);
    std::set<unsigned> keys;
    keysOf(argv[1], keys);
    keysOf(argv[2], keys);
    FILE* out = std::fopen(argv[3], "w");
    if (!out) return 3;
    for (unsigned d : keys) {
        const std::vector<unsigned>& v = m.getConnectedModuleDetIds(d);
        std::fprintf(out, "%u %zu", d, v.size());
        for (unsigned c : v) std::fprintf(out, " %u", c);
        std::fprintf(out, "\n");
    }
    std::fclose(out);
    return 0;
}
