// Op-counting build of the CPU oracle -- measurement infrastructure only.
//
// Compiles oracle/sdl_oracle.cc with float and double replaced by the counting
// types of opcount.h, runs the chain on generated events stage by stage, and
// prints the algorithmic work of each stage: the floating-point operations
// along the path every candidate actually takes (early exits included), as the
// reference's loops evaluate them (Event.cc:245-1062).  The object counts are
// printed too, so a run can be checked against the plain oracle.
//
// Usage: oracle_opcount <workload> <geomdir> <event0> <nevents> [libm]
// Output: one JSON object per event on stdout.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "opcount.h"
#include "sdl_gen.h"
#include "sdl_oracle.h"

unsigned long long opc::ops = 0;

#define ORACLE_OPCOUNT 1
#define float opc::F32
#define double opc::F64
#include "sdl_oracle.cc"
#undef float
#undef double

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s <workload> <geomdir> <event0> <nevents> [libm]\n", argv[0]);
        return 2;
    }
    const int workload = std::atoi(argv[1]);
    const std::string dir = argv[2];
    const int e0 = std::atoi(argv[3]), ne = std::atoi(argv[4]);
    const int libm = argc > 5 ? std::atoi(argv[5]) : ORACLE_LIBM_CR;
    const int kind = sdlgen::geometryKindOf(workload);
    sdlgen::Geometry gg;
    sdlgen::buildGeometry(kind, gg, kind == 1 ? 50000 : 200000);
    const std::string t = dir + "/opc_tilted_" + std::to_string(kind) + ".txt";
    const std::string e = dir + "/opc_endcap_" + std::to_string(kind) + ".txt";
    const std::string m = dir + "/opc_map_" + std::to_string(kind) + ".txt";
    sdlgen::writeGeometryFiles(gg, t, e, m);
    oracle::Geometry g;
    if (!oracle::loadGeometry(g, t.c_str(), e.c_str(), m.c_str())) return 1;
    const sdlgen::EventConfig cfg = sdlgen::presetConfig(workload);
    for (int ie = e0; ie < e0 + ne; ++ie) {
        sdlgen::Event gev;
        sdlgen::generateEvent(gg, cfg, uint64_t(ie), gev);
        const size_t n = gev.x.size();
        std::vector<opc::F32> x(n), y(n), z(n);
        for (size_t i = 0; i < n; ++i) x[i] = gev.x[i], y[i] = gev.y[i], z[i] = gev.z[i];
        oracle::Event ev;
        opc::ops = 0;
        oracle::fillEvent(ev, &g, libm, int(n), x.data(), y.data(), z.data(), gev.detId.data(), gev.idx.data());
        unsigned long long f[7];
        f[0] = opc::ops;
        opc::ops = 0; ev.createMiniDoublets();    f[1] = opc::ops;
        opc::ops = 0; ev.createSegments();        f[2] = opc::ops;
        opc::ops = 0; ev.createTriplets();        f[3] = opc::ops;
        opc::ops = 0; ev.createTracklets();       f[4] = opc::ops;
        opc::ops = 0; ev.createTCFromTriplets();  f[5] = opc::ops;
        opc::ops = 0; ev.createTCFromTracklets(); f[6] = opc::ops;
        size_t ntcp = 0, ntcl = 0;
        for (auto& tc : ev.tcs) (tc.first == 0 ? ntcp : ntcl)++;
        std::printf("{\"workload\": %d, \"event\": %d, \"hits\": %zu, "
                    "\"flops\": {\"hits\": %llu, \"md\": %llu, \"sg\": %llu, \"tp\": %llu, \"tl\": %llu, \"tcp\": %llu, \"tcl\": %llu}, "
                    "\"candidates\": [%llu, %llu, %llu, %llu, %llu], "
                    "\"objects\": [%zu, %zu, %zu, %zu, %zu, %zu]}\n",
                    workload, ie, n, f[0], f[1], f[2], f[3], f[4], f[5], f[6], ev.nCand[0], ev.nCand[1], ev.nCand[2],
                    ev.nCand[3], ev.nCand[4], ev.mds.size(), ev.sgs.size(), ev.tps.size(), ev.tls.size(), ntcp, ntcl);
        std::fflush(stdout);
    }
    return 0;
}
