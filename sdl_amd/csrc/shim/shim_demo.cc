// A host program written against the SDL::Event API (as a user of the reference
// would), running one generated event through the MI355X library.
// Usage: sdl_shim_demo <workload> <event> <geomdir> [layers|pseudo]
// Runs the module-map chain, with "layers" the layer-pair chain
// (createMiniDoublets, createSegments, createTracklets, createTrackCandidates),
// with "pseudo" the module-map chain on pseudo mini-doublets;
// prints the counters, a checksum of every object set, a checksum of the MD, SG,
// TL and TC-from-tracklets objects joined with their per-object values, and the
// value column names, one line each.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "SDL/Event.h"
#include "sdl_gen.h"

template <class V>
unsigned long long checksum(const V& v) {
    unsigned long long s = 0;
    for (const auto& a : v) {
        unsigned long long h = 1469598103934665603ULL;
        for (int x : a) h = (h ^ (unsigned)x) * 1099511628211ULL;
        s += h;  // order-independent
    }
    return s;
}

// Order-independent checksum of (object tuple, value bits) rows.
template <class O, class V>
unsigned long long checksumWithValues(const O& objs, const V& vals) {
    if (objs.size() != vals.size()) return 0;
    unsigned long long s = 0;
    for (size_t i = 0; i < objs.size(); ++i) {
        unsigned long long h = 1469598103934665603ULL;
        for (int x : objs[i]) h = (h ^ (unsigned)x) * 1099511628211ULL;
        for (float f : vals[i]) {
            unsigned bits = 0x7fc00000u;  // every NaN as the canonical quiet NaN
            if (f == f) std::memcpy(&bits, &f, 4);
            h = (h ^ bits) * 1099511628211ULL;
        }
        s += h;
    }
    return s;
}

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    const int workload = std::atoi(argv[1]);
    const int ievent = std::atoi(argv[2]);
    const std::string dir = argv[3];
    const int kind = (workload == 1 || workload == 4) ? 1 : 2;
    sdlgen::Geometry g;
    sdlgen::buildGeometry(kind, g, kind == 1 ? 50000 : 200000);
    const std::string t = dir + "/shim_tilted.txt", e = dir + "/shim_endcap.txt", m = dir + "/shim_map.txt";
    sdlgen::writeGeometryFiles(g, t, e, m);
    SDL::tiltedGeometry.load(t);
    SDL::endcapGeometry.load(e);
    SDL::moduleConnectionMap.load(m);
    sdlgen::Event ev;
    sdlgen::generateEvent(g, sdlgen::presetConfig(workload), ievent, ev);

    SDL::Event event;
    for (size_t i = 0; i < ev.x.size(); ++i) event.addHitToModule(SDL::Hit(ev.x[i], ev.y[i], ev.z[i], ev.idx[i]), ev.detId[i]);
    const bool layers = argc > 4 && std::string(argv[4]) == "layers";
    const bool pseudo = argc > 4 && std::string(argv[4]) == "pseudo";
    if (pseudo) event.createPseudoMiniDoubletsFromAnchorModule();
    else event.createMiniDoublets();
    if (layers) {
        event.createSegments();
        event.createTracklets();
        event.createTrackCandidates();
    } else {
        event.createSegmentsWithModuleMap();
        event.createTriplets();
        event.createTrackletsWithModuleMap();
        event.createTrackCandidatesFromTriplets();
        event.createTrackCandidatesFromTracklets();
    }
    std::printf("counts %u %u %u %u %u %u %u %u %u %u\n", event.getNumberOfMiniDoubletCandidates(),
                event.getNumberOfSegmentCandidates(), event.getNumberOfTripletCandidates(),
                event.getNumberOfTrackletCandidates(), event.getNumberOfTrackCandidateCandidates(),
                event.getNumberOfMiniDoublets(), event.getNumberOfSegments(), event.getNumberOfTriplets(),
                event.getNumberOfTracklets(), event.getNumberOfTrackCandidates());
    std::printf("checksums %llu %llu %llu %llu %llu %llu\n", checksum(event.getMiniDoublets()),
                checksum(event.getSegments()), checksum(event.getTriplets()), checksum(event.getTracklets()),
                checksum(event.getTrackCandidatesFromTriplets()), checksum(event.getTrackCandidatesFromTracklets()));
    std::printf("valuesums %llu %llu %llu %llu\n",
                checksumWithValues(event.getMiniDoublets(), event.getMiniDoubletVariables()),
                checksumWithValues(event.getSegments(), event.getSegmentVariables()),
                checksumWithValues(event.getTracklets(), event.getTrackletVariables()),
                checksumWithValues(event.getTrackCandidatesFromTracklets(),
                                   event.getTrackCandidatesFromTrackletsVariables()));
    std::string names;
    for (const char* t : {"MD", "SG", "TL", "TC"})
        for (const auto& n : SDL::Event::variableNames(t)) names += std::string(names.empty() ? "" : ",") + t + "." + n;
    std::printf("valuenames %s\n", names.c_str());
    return 0;
}
