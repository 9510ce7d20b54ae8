// Test infrastructure: the module detId decoding of the compiled reference
// (SDL::Module::parse*, Module.cc:159-358, in oracle/_ref/libsdl_ref.so)
// against the oracle's restatement (oracle_module_decode) and the library's
// (sdl_module_decode, the decoding behind every device table and the shim's
// SDL::Module), field by field.
//
// Every detId whose bits the decoding reads -- 0, 2-8, 10-16 and 18-27, all 2^25
// combinations -- once with the other bits (1, 9, 17, 28-31) clear and once with
// them set from a hash of the detId: the oracle and the library on all of them,
// the reference on all the detIds it can categorise and on one in 1024 of those
// it cannot (where it prints a warning on every call: subdet not 4 or 5, endcap
// side not 1 or 2, barrel side 0, barrel side 1 or 2 beyond layer 3); plus
// closed-form checks on the sensors of
// real modules (bits 1:0 = 01 or 10, subdet 4 or 5): the partner of the partner
// is the sensor itself, the partners differ only in bits 1:0, exactly one of
// them is the lower sensor, both have the same module type and inversion, and
// of a PS module's two sensors exactly one is the pixel sensor.
//
// Usage: decode_check [threads]   prints "ok <detIds checked>" or the first mismatch.
#include <atomic>
#include <fcntl.h>
#include <unistd.h>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <thread>
#include <vector>

#include "Module.h"

extern "C" int sdl_module_decode(unsigned int detId, unsigned int out[11]);
extern "C" void oracle_module_decode(unsigned detId, unsigned out[11]);

namespace {

void reference(unsigned d, unsigned v[11]) {
    v[0] = SDL::Module::parseSubdet(d);
    v[1] = SDL::Module::parseSide(d);
    v[2] = SDL::Module::parseLayer(d);
    v[3] = SDL::Module::parseRod(d);
    v[4] = SDL::Module::parseRing(d);
    v[5] = SDL::Module::parseModule(d);
    v[6] = SDL::Module::parseIsLower(d);
    v[7] = SDL::Module::parseIsInverted(d);
    v[8] = SDL::Module::parsePartnerDetId(d);
    v[9] = SDL::Module::parseModuleType(d) == SDL::Module::TwoS;
    v[10] = SDL::Module::parseModuleLayerType(d) == SDL::Module::Strip;
}

// Spreads the 25 read bits of c over the detId positions that hold them.
unsigned spread(unsigned c) {
    return (c & 1u) | (((c >> 1) & 127u) << 2) | (((c >> 8) & 127u) << 10) | (((c >> 15) & 1023u) << 18);
}
const unsigned kUnread = (1u << 1) | (1u << 9) | (1u << 17) | 0xF0000000u;

unsigned hash(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    return x ^ (x >> 16);
}

std::atomic<bool> g_failed{false};

bool same(const unsigned* a, const unsigned* b) {
    for (int k = 0; k < 11; ++k)
        if (a[k] != b[k]) return false;
    return true;
}

void report(const char* what, unsigned d, const unsigned* a, const unsigned* b) {
    if (g_failed.exchange(true)) return;
    std::fprintf(stderr, "mismatch (%s) at detId %u:", what, d);
    for (int k = 0; k < 11; ++k) std::fprintf(stderr, " %u/%u", a[k], b[k]);
    std::fprintf(stderr, "\n");
}

// The detIds for which the reference's parseIsInverted cannot categorise the module.
bool uncategorised(unsigned d) {
    const unsigned sub = (d >> 25) & 7u;
    if (sub == 4u) {
        const unsigned side = (d >> 23) & 3u;
        return side != 1u && side != 2u;
    }
    if (sub == 5u) {
        const unsigned side = (d >> 18) & 3u, layer = (d >> 20) & 7u;
        return side == 0u || (side != 3u && layer >= 4u);
    }
    return true;
}

void checkOne(unsigned d, bool withReference) {
    unsigned r[11], o[11], l[11];
    oracle_module_decode(d, o);
    if (sdl_module_decode(d, l) != 0) { report("library error", d, o, o); return; }
    if (!same(o, l)) report("library vs oracle", d, o, l);
    if (!withReference) return;
    reference(d, r);
    if (!same(r, o)) report("oracle", d, r, o);
    if (!same(r, l)) report("library", d, r, l);
    const unsigned low = d & 3u, sub = r[0];
    if ((low == 1u || low == 2u) && (sub == 4u || sub == 5u) && !uncategorised(d)) {
        unsigned p[11];
        reference(r[8], p);
        const bool ok = p[8] == d && ((r[8] ^ d) & ~3u) == 0 && p[6] != r[6] && p[7] == r[7] && p[9] == r[9] &&
                        (r[9] == 1u ? (p[10] == 1u && r[10] == 1u) : (p[10] != r[10]));
        if (!ok) report("partner", d, r, p);
    }
}

}  // namespace

int main(int argc, char** argv) {
    const int threads = argc > 1 ? std::atoi(argv[1]) : 8;
    // the reference's categorisation warnings (SDL::cout, on standard output)
    std::fflush(stdout);
    const int devnull = open("/dev/null", O_WRONLY);
    if (devnull < 0 || dup2(devnull, 1) < 0) return 2;
    const unsigned n = 1u << 25;
    std::atomic<unsigned> next{0};
    auto work = [&]() {
        for (;;) {
            const unsigned c0 = next.fetch_add(65536);
            if (c0 >= n || g_failed) return;
            for (unsigned c = c0; c < c0 + 65536; ++c) {
                const unsigned d = spread(c), h = hash(c);
                const bool sampled = (h >> 4) % 1024u == 0u;
                checkOne(d, !uncategorised(d) || sampled);
                const unsigned d2 = d | (h & kUnread);
                checkOne(d2, !uncategorised(d2) || sampled);
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t) th.emplace_back(work);
    for (auto& t : th) t.join();
    if (g_failed) return 1;
    std::fprintf(stderr, "ok %llu\n", 2ull * n);
    return 0;
}
