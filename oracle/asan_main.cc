// Test infrastructure: the oracle and the event generator built with
// AddressSanitizer and UndefinedBehaviorSanitizer (g++ -fsanitize=address,
// undefined, no recovery; oracle/Makefile target `asan`), run over small
// events through every stage sequence the oracle restates: the module-map
// chain (both libm policies), the layer-pair chain, navigation tracklets,
// pseudo mini-doublets, DefaultNm1 and AllComb algorithms, per-object values
// and the candidate explanation.  A sanitizer report ends the process with a
// non-zero status; a clean run prints "ok <objects>".
//
// Usage: oracle_asan <geometry dir>   (the three text files are written there)
#include <cstdio>
#include <string>
#include <vector>

#include "sdl_gen.h"
#include "sdl_oracle.h"

namespace {

long long runOne(void* g, const sdlgen::Event& ev, int libm, const std::vector<int>& seq) {
    void* r = seq.empty() ? oracle_run(g, int(ev.x.size()), ev.x.data(), ev.y.data(), ev.z.data(), ev.detId.data(),
                                       ev.idx.data(), libm, 6)
                          : oracle_run_sequence(g, int(ev.x.size()), ev.x.data(), ev.y.data(), ev.z.data(),
                                                ev.detId.data(), ev.idx.data(), libm, seq.data(), int(seq.size()));
    if (!r) return -1;
    long long n = 0;
    for (int s = 0; s < 6; ++s) {
        const long c = oracle_result_count(r, s);
        static const int width[6] = {2, 4, 8, 8, 16, 16};
        std::vector<int> t(size_t(c) * width[s] + 1);
        oracle_result_copy(r, s, t.data());
        const int nv = oracle_var_count(s);
        if (nv > 0) {
            std::vector<float> v(size_t(c) * nv + 1);
            oracle_result_vars(r, s, v.data());
        }
        n += c;
    }
    unsigned long long cnt[10], lay[110];
    oracle_result_counters(r, cnt);
    oracle_result_layers(r, lay);
    oracle_result_free(r);
    return n;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s geometry-dir\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    long long total = 0;
    // (workload, events, sequences): workload 1 the flat barrel, 12 and 15 the
    // full detector at 300 tracks (plain and hard mix), 11 the flat barrel at a
    // few tracks (AllComb sequences stay small)
    struct Case {
        int workload, events;
        std::vector<std::vector<int>> seqs;
    };
    const std::vector<Case> cases = {
        {1, 2, {{}, {0, 1, 2, 103, 4, 5}, {0, 6, 7, 8}, {0, 1, 9, 5}, {10, 0, 1, 2, 3, 4, 5}}},
        {12, 1, {{}, {0, 6, 107, 8}, {0, 1, 109, 5}}},
        {15, 1, {{}}},
        {11, 1, {{100, 1, 2, 3, 4, 5}, {0, 101, 2, 3, 4, 5}, {0, 1, 2, 203, 4, 5}, {0, 1, 2, 3, 104, 105}}},
    };
    for (const Case& c : cases) {
        sdlgen::Geometry geo;
        const int kind = sdlgen::geometryKindOf(c.workload);
        sdlgen::buildGeometry(kind, geo, kind == 1 ? 50000 : 200000);
        const std::string t = dir + "/tilted_" + std::to_string(kind) + ".txt";
        const std::string e = dir + "/endcap_" + std::to_string(kind) + ".txt";
        const std::string m = dir + "/map_" + std::to_string(kind) + ".txt";
        sdlgen::writeGeometryFiles(geo, t, e, m);
        void* g = oracle_geom_load(t.c_str(), e.c_str(), m.c_str());
        if (!g) {
            std::fprintf(stderr, "geometry load failed\n");
            return 3;
        }
        const sdlgen::EventConfig cfg = sdlgen::presetConfig(c.workload);
        for (int ie = 0; ie < c.events; ++ie) {
            sdlgen::Event ev;
            sdlgen::generateEvent(geo, cfg, ie, ev);
            for (const auto& seq : c.seqs)
                for (int libm : {ORACLE_LIBM_GLIBC, ORACLE_LIBM_CR}) {
                    const long long n = runOne(g, ev, libm, seq);
                    if (n < 0) {
                        std::fprintf(stderr, "run failed: workload %d event %d\n", c.workload, ie);
                        return 4;
                    }
                    total += n;
                }
            // the candidate explanation of the event's first mini-doublet
            void* r = oracle_run(g, int(ev.x.size()), ev.x.data(), ev.y.data(), ev.z.data(), ev.detId.data(),
                                 ev.idx.data(), ORACLE_LIBM_GLIBC, 1);
            const long nmd = r ? oracle_result_count(r, 0) : 0;
            if (nmd > 0) {
                std::vector<int> md(size_t(nmd) * 2);
                oracle_result_copy(r, 0, md.data());
                const int b[4] = {0, 0, 0, 0};
                double out[9];
                oracle_explain(g, int(ev.x.size()), ev.x.data(), ev.y.data(), ev.z.data(), ev.detId.data(),
                               ev.idx.data(), 0, md.data(), b, ORACLE_LIBM_CR, out);
            }
            if (r) oracle_result_free(r);
        }
        oracle_geom_free(g);
    }
    std::fprintf(stderr, "ok %lld\n", total);
    return 0;
}
