// Test support: the host-only entry points of the C-ABI (geometry text parsing
// and the binary cache, detId decoding, the connection-map builder) run in a
// process built with AddressSanitizer and UndefinedBehaviorSanitizer on the
// host code (the library's device code is built as usual).  No GPU is used:
// these entry points never touch one.  A sanitizer report ends the process
// with a non-zero status; a clean run prints "ok".
//
// Usage: sdl_host_asan <scratch dir>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../../include/sdl_amd.h"
#include "../gen/sdl_gen.h"

namespace {

int fail(const char* what) {
    std::fprintf(stderr, "failed: %s (%s)\n", what, sdl_last_error());
    return 1;
}

std::vector<char> readAll(const std::string& p) {
    std::ifstream f(p, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}
void writeAll(const std::string& p, const std::vector<char>& b) {
    std::ofstream f(p, std::ios::binary);
    f.write(b.data(), std::streamsize(b.size()));
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    for (int kind : {1, 2}) {
        sdlgen::Geometry geo;
        sdlgen::buildGeometry(kind, geo, 20000);
        const std::string t = dir + "/t.txt", e = dir + "/e.txt", m = dir + "/m.txt", b = dir + "/g.bin";
        sdlgen::writeGeometryFiles(geo, t, e, m);
        if (sdl_geom_convert(t.c_str(), e.c_str(), m.c_str(), b.c_str()) != 0) return fail("convert");
        int nm = 0;
        long long nc = 0;
        if (sdl_geom_check_binary(b.c_str(), &nm, &nc) != 0 || nm <= 0) return fail("check");
        // damaged caches are rejected: flipped bytes, truncation, extension
        const std::vector<char> good = readAll(b);
        std::mt19937 rng(kind);
        for (int k = 0; k < 64; ++k) {
            std::vector<char> bad = good;
            bad[rng() % bad.size()] ^= char(1 + rng() % 255);
            writeAll(b, bad);
            if (sdl_geom_check_binary(b.c_str(), &nm, &nc) == 0) return fail("a flipped byte was accepted");
        }
        for (size_t cut : {size_t(0), size_t(7), good.size() / 2, good.size() - 1}) {
            writeAll(b, std::vector<char>(good.begin(), good.begin() + long(cut)));
            if (sdl_geom_check_binary(b.c_str(), &nm, &nc) == 0) return fail("a truncated cache was accepted");
        }
        std::vector<char> longer = good;
        longer.push_back(0);
        writeAll(b, longer);
        if (sdl_geom_check_binary(b.c_str(), &nm, &nc) == 0) return fail("an extended cache was accepted");
        // malformed text files: missing columns, garbage, a duplicated connection
        {
            std::ofstream f(dir + "/bad_map.txt");
            f << "123 3 5 6\n456 2 7 7\nxyz\n";
        }
        sdl_geom_convert(t.c_str(), e.c_str(), (dir + "/bad_map.txt").c_str(), b.c_str());
        // connection map from random tracks, written and then merged
        std::vector<long long> off{0};
        std::vector<int> station;
        std::vector<unsigned> det;
        std::uniform_int_distribution<size_t> pick(0, geo.sensors.size() - 1);
        for (int tr = 0; tr < 2000; ++tr) {
            const int n = 1 + int(rng() % 12);
            for (int k = 0; k < n; ++k) {
                station.push_back(k / 2);
                det.push_back(geo.sensors[pick(rng)].detId);
            }
            off.push_back(off.back() + n);
        }
        const std::string mp = dir + "/built.txt";
        if (sdl_map_build(2000, off.data(), station.data(), det.data(), mp.c_str(), 0) != 0) return fail("map build");
        if (sdl_map_build(1000, off.data(), station.data(), det.data(), mp.c_str(), 1) != 0) return fail("map merge");
    }
    unsigned out[11];
    for (unsigned c = 0; c < (1u << 22); ++c) sdl_module_decode(c * 977u + (c << 9), out);
    std::fprintf(stderr, "ok\n");
    return 0;
}
