// Test infrastructure: single functions of the compiled reference
// (oracle/_ref/libsdl_ref.so) against the oracle's restatements
// (oracle_kat_* in oracle/_build/liboracle.so), value for value:
//
//  * MathUtil::Phi_mpi_pi (MathUtil.cc:3-19) on every float of [-1000, 1000]
//    whose low 8 mantissa bits are zero, plus +-0, NaN and the floats next to
//    +-pi, +-2pi and +-3pi (not on +-inf, where the reference's loop never
//    ends);
//  * MathUtil::ATan2 (MathUtil.cc:21-27) on a grid of signed values with zeros;
//  * MathUtil::getCenterFromThreePoints (MathUtil.cc:44-106) on random triples
//    and on the degenerate ones its branches single out (A and B, or B and C,
//    on a vertical or horizontal line; collinear points; coincident points);
//  * MiniDoublet::isTighterTiltedModules and moduleGapSize
//    (MiniDoublet.cc:1242-1332) on every sensor of every barrel layer, side and
//    rod, and of every endcap side, layer and ring (module numbers 0-127);
//  * MiniDoublet::shiftStripHits (MiniDoublet.cc:989-1209) on random hit pairs
//    on every lower module of the tilted-barrel and endcap geometry files given
//    (whose strip slopes include the reference's SDL_INF and zero), with the
//    same files loaded into the reference's geometry singletons.
//
// Usage: kat_check <tilted> <endcap> <map>   prints "ok <values checked>" on
// stderr or the first difference (exit status 1).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "EndcapGeometry.h"
#include "Hit.h"
#include "MathUtil.h"
#include "MiniDoublet.h"
#include "Module.h"
#include "TiltedGeometry.h"

extern "C" {
float oracle_kat_phi_mpi_pi(float x);
float oracle_kat_atan2(float y, float x);
void oracle_kat_center(const float* x, const float* y, float* out2);
float oracle_kat_module_gap(unsigned detId, int* tighter);
void oracle_kat_shift(void* geom, unsigned lowerDetId, const float* lo, const float* up, float* out3);
void* oracle_geom_load(const char* tiltedPath, const char* endcapPath, const char* mapPath);
}

namespace {

long long g_checked = 0;

uint32_t bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
bool same(float a, float b) { return (std::isnan(a) && std::isnan(b)) || bits(a) == bits(b); }

bool expect(const char* what, float ref, float orc, const std::string& where) {
    ++g_checked;
    if (same(ref, orc)) return true;
    std::fprintf(stderr, "%s differs at %s: reference %.9g (0x%08x), oracle %.9g (0x%08x)\n", what, where.c_str(),
                 ref, bits(ref), orc, bits(orc));
    return false;
}

std::string fmt(const char* f, double a, double b = 0, double c = 0) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), f, a, b, c);
    return buf;
}

bool checkPhi() {
    std::vector<float> xs = {0.f, -0.f, NAN};
    const float anchors[] = {float(M_PI), float(2 * M_PI), float(3 * M_PI)};
    for (float a : anchors)
        for (float s : {1.f, -1.f})
            for (int k = -4; k <= 4; ++k) {
                float v = s * a;
                for (int j = 0; j < std::abs(k); ++j) v = std::nextafter(v, k > 0 ? INFINITY : -INFINITY);
                xs.push_back(v);
            }
    for (uint32_t u = 0; u < 0x80000000u; u += 256) {
        float f;
        std::memcpy(&f, &u, 4);
        if (!(f <= 1000.f)) break;
        xs.push_back(f);
        xs.push_back(-f);
    }
    for (float x : xs)
        if (!expect("Phi_mpi_pi", SDL::MathUtil::Phi_mpi_pi(x), oracle_kat_phi_mpi_pi(x), fmt("x = %.9g", x)))
            return false;
    return true;
}

bool checkATan2() {
    std::vector<float> v = {0.f, -0.f, 1e-30f, -1e-30f, 1e-3f, -1e-3f, 0.5f, -0.5f, 1.f, -1.f, 3.f, -3.f, 1e4f, -1e4f};
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> u(-200.f, 200.f);
    for (int k = 0; k < 200; ++k) v.push_back(u(rng));
    for (float y : v)
        for (float x : v)
            if (!expect("ATan2", SDL::MathUtil::ATan2(y, x), oracle_kat_atan2(y, x), fmt("y = %.9g, x = %.9g", y, x)))
                return false;
    return true;
}

bool checkCenter() {
    std::mt19937 rng(11);
    std::uniform_real_distribution<float> u(-120.f, 120.f);
    std::vector<std::array<float, 6>> cases;
    for (int k = 0; k < 100000; ++k) cases.push_back({u(rng), u(rng), u(rng), u(rng), u(rng), u(rng)});
    for (int k = 0; k < 2000; ++k) {
        const float a = u(rng), b = u(rng), c = u(rng), d = u(rng), e = u(rng), f = u(rng);
        cases.push_back({a, b, a, d, e, f});   // A, B on a vertical line
        cases.push_back({a, b, c, d, c, f});   // B, C on a vertical line
        cases.push_back({a, b, c, b, e, f});   // A, B on a horizontal line
        cases.push_back({a, b, c, d, c, d});   // B = C
        cases.push_back({a, b, a, b, a, b});   // all three coincide
        const float t1 = 0.25f * (u(rng) + 120.f) / 240.f, t2 = 2.f * t1;
        cases.push_back({a, b, a + t1 * (c - a), b + t1 * (d - b), a + t2 * (c - a), b + t2 * (d - b)});  // collinear
        cases.push_back({a, b, a + 1.f, b + 2.f, a + 2.f, b + 4.f});  // collinear, exact slopes
    }
    for (const auto& p : cases) {
        SDL::Hit A(p[0], p[1], 0.f), B(p[2], p[3], 0.f), C(p[4], p[5], 0.f);
        const SDL::Hit c = SDL::MathUtil::getCenterFromThreePoints(A, B, C);
        const float x[3] = {p[0], p[2], p[4]}, y[3] = {p[1], p[3], p[5]};
        float o[2];
        oracle_kat_center(x, y, o);
        const std::string w = fmt("A (%.9g, %.9g) ...", p[0], p[1]);
        if (!expect("center x", c.x(), o[0], w) || !expect("center y", c.y(), o[1], w)) return false;
    }
    return true;
}

bool checkGap(std::vector<unsigned>& sensors) {
    // barrel: subdet 5, side 3 (center) layers 1-6, sides 1-2 (tilted) layers 1-3
    for (unsigned side = 1; side <= 3; ++side)
        for (unsigned layer = 1; layer <= (side == 3 ? 6u : 3u); ++layer)
            for (unsigned rod = 0; rod < 128; ++rod)
                for (unsigned mod = 0; mod < 128; ++mod)
                    for (unsigned b = 1; b <= 2; ++b)
                        sensors.push_back((5u << 25) | (layer << 20) | (side << 18) | (rod << 10) | (mod << 2) | b);
    // endcap: subdet 4, sides 1-2, layers 1-5, rings 1-15
    for (unsigned side = 1; side <= 2; ++side)
        for (unsigned layer = 1; layer <= 5; ++layer)
            for (unsigned ring = 1; ring <= 15; ++ring)
                for (unsigned mod = 0; mod < 128; ++mod)
                    for (unsigned b = 1; b <= 2; ++b)
                        sensors.push_back((4u << 25) | (side << 23) | (layer << 18) | (ring << 12) | (mod << 2) | b);
    for (unsigned d : sensors) {
        SDL::Module m(d);
        int tighter = -1;
        const float gap = oracle_kat_module_gap(d, &tighter);
        const std::string w = fmt("detId %.0f", double(d));
        if (!expect("moduleGapSize", SDL::MiniDoublet::moduleGapSize(m), gap, w)) return false;
        if (!expect("isTighterTiltedModules", float(SDL::MiniDoublet::isTighterTiltedModules(m)), float(tighter), w))
            return false;
    }
    return true;
}

// detIds in the first column of a geometry file
std::vector<unsigned> detIdsOf(const char* path) {
    std::vector<unsigned> out;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        unsigned d;
        if (is >> d) out.push_back(d);
    }
    return out;
}

bool checkShift(const char* tilted, const char* endcap, const char* map) {
    SDL::tiltedGeometry.load(tilted);
    SDL::endcapGeometry.load(endcap);
    void* g = oracle_geom_load(tilted, endcap, map);
    if (!g) {
        std::fprintf(stderr, "oracle could not load the geometry\n");
        return false;
    }
    std::vector<unsigned> ids = detIdsOf(tilted);
    const std::vector<unsigned> ec = detIdsOf(endcap);
    ids.insert(ids.end(), ec.begin(), ec.end());
    std::mt19937 rng(5);
    std::uniform_real_distribution<float> ang(-3.14159f, 3.14159f), rr(20.f, 110.f), zz(-260.f, 260.f), d(-0.4f, 0.4f);
    int modules = 0;
    for (unsigned id : ids) {
        SDL::Module m(id);
        if (!m.isLower()) continue;
        const bool endcapMod = m.subdet() == SDL::Module::Endcap;
        if (!endcapMod && m.side() == SDL::Module::Center) continue;  // shiftStripHits runs on tilted and endcap modules
        ++modules;
        for (int k = 0; k < 24; ++k) {
            const float phi = ang(rng), r = rr(rng), z0 = zz(rng);
            const float lo[3] = {r * std::cos(phi), r * std::sin(phi), z0 == 0.f ? 1.f : z0};
            const float up[3] = {lo[0] + d(rng), lo[1] + d(rng), lo[2] + d(rng)};
            SDL::Hit hl(lo[0], lo[1], lo[2]), hu(up[0], up[1], up[2]);
            const auto t = SDL::MiniDoublet::shiftStripHits(hl, hu, m);
            float o[3];
            oracle_kat_shift(g, id, lo, up, o);
            const std::string w = fmt("lower module %.0f, hit (%.9g, %.9g, ...)", double(id), lo[0], lo[1]);
            if (!expect("shifted x", std::get<0>(t), o[0], w) || !expect("shifted y", std::get<1>(t), o[1], w) ||
                !expect("shifted z", std::get<2>(t), o[2], w))
                return false;
        }
    }
    if (modules == 0) {
        std::fprintf(stderr, "no tilted or endcap lower modules in the geometry files\n");
        return false;
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s tilted endcap map\n", argv[0]);
        return 2;
    }
    std::vector<unsigned> sensors;
    if (!checkPhi()) return 1;
    std::fprintf(stderr, "Phi_mpi_pi: %lld values\n", g_checked);
    if (!checkATan2()) return 1;
    std::fprintf(stderr, "ATan2: %lld values\n", g_checked);
    if (!checkCenter()) return 1;
    std::fprintf(stderr, "getCenterFromThreePoints: %lld values\n", g_checked);
    if (!checkGap(sensors)) return 1;
    std::fprintf(stderr, "moduleGapSize, isTighterTiltedModules: %lld values\n", g_checked);
    if (!checkShift(argv[1], argv[2], argv[3])) return 1;
    std::fprintf(stderr, "ok %lld\n", g_checked);
    return 0;
}
