// Synthetic outer-tracker detector and seeded event generator for the SDL chain.
//
// The reference ships no geometry, connection map or events (its host driver
// lived elsewhere), so the workloads are generated here:
//
//  * a detector whose module detIds follow the reference's 29-bit layout and
//    lower/upper parity rules (Module.cc:159-358), with tilted-geometry lines
//    "detid drdz slope" (TiltedGeometry.cc:36), endcap-geometry lines of nine
//    columns whose 7th column is read as the centroid phi
//    (EndcapGeometry.cc:48,57-58) and connection-map lines "detid n d1..dn"
//    (ModuleConnectionMap.cc:36-44) built by propagating straight-from-the-IP
//    helices through the detector;
//  * events of charged helices in a 3.8 T solenoid field with Gaussian hit
//    smearing (pitch/sqrt(12)), optional multiple-scattering kicks and
//    uniform noise hits, drawn from a counter-based generator keyed by
//    (workload, event, track) so that every event regenerates identically on
//    any machine with the same libm.
//
// Everything is plain host C++; nothing here runs on the GPU.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sdlgen {

enum GeomKind {
    GEOM_BARREL_FLAT = 1,  // six flat barrel layers, |z| < 120 cm
    GEOM_FULL = 2          // flat + tilted inner barrel and 2x5 endcap disks
};

struct Sensor {
    uint32_t detId;
    double c[3];      // centre
    double n[3];      // unit normal (points away from the interaction point)
    double u[3];      // precise in-plane direction (r-phi like)
    double v[3];      // coarse in-plane direction (strip length)
    double hu, hv;    // half extents along u and v
    double sig_u, sig_v;  // Gaussian smearing widths
    int module;       // owning module index
};

struct Module {
    int sensor[2];    // [0] = lower (closer to the IP), [1] = upper
    int surf;         // index into Geometry::surfs
    bool barrel, ps, tilted;
    int layer;
    double drdz;      // |dr/dz| of the module plane (tilted barrel); 0 for flat
    double slope;     // x-y slope of the module measurement direction (123456789 = vertical)
    double centroidPhi, centroidR, centroidZ;
};

struct Ring {
    double pos, half;     // z centre / half extent (barrel) or r centre / half extent (endcap)
    int nphi;
    double phi0;
    std::vector<int> modules;  // module index by phi slot
};

struct Surface {
    bool barrel;
    double rz;        // cylinder radius (barrel) or signed disk z (endcap)
    int layer;        // 1..6 barrel, 1..5 endcap
    std::vector<Ring> rings;
};

struct Geometry {
    int kind = 0;
    std::vector<Sensor> sensors;
    std::vector<Module> modules;
    std::vector<Surface> surfs;
    // connection map: lower detId -> sorted unique lower detIds of the next layer
    std::vector<std::pair<uint32_t, std::vector<uint32_t>>> conn;
};

struct EventConfig {
    int workload = 1;          // 1..5, the workload family the event belongs to
    int ntracks_min = 500, ntracks_max = 500;
    int n_hard = 0;            // extra tracks with pT ~ U[5,100] GeV
    int pt_mode = 0;           // 0: 1/pT ~ U[invpt_lo, invpt_hi]; 1: p(pT) ~ pT exp(-pT/0.6), pT > 0.8
    double invpt_lo = 0.1, invpt_hi = 1.0;
    double eta_max = 0.8;
    double sigma_z = 5.0;
    double noise_rate = 0.05;  // mean noise hits per sensor
    bool ms = false;           // multiple-scattering kicks
    double frac_lowpt = 0.0, lowpt_lo = 0.5, lowpt_hi = 0.8;
    double frac_displaced = 0.0, rv_max = 30.0;
    uint64_t seed = 0x5d1c0ffee;
};

struct Event {
    std::vector<float> x, y, z;
    std::vector<uint32_t> detId;
    std::vector<int> idx;
    std::vector<int> simTrack;   // -1 for noise
    int ntracks = 0;
};

// Builds the detector (deterministic, no randomness) including the connection map.
void buildGeometry(int kind, Geometry& g, int mapTracks = 200000);
// Geometry kind of a workload: the flat six-layer barrel (1, 4, 11) or the full detector.
int geometryKindOf(int workload);

// Writes the three text files the reference loads.
void writeGeometryFiles(const Geometry& g, const std::string& tiltedPath,
                        const std::string& endcapPath, const std::string& mapPath);

// Returns a preset for workload 1..5 (the five benchmark configurations) or
// 12 / 15 (300-track slices of workloads 2 / 5 used by the parity tests).
EventConfig presetConfig(int workload);

// Generates event number `ievent` of the given configuration.
void generateEvent(const Geometry& g, const EventConfig& cfg, uint64_t ievent, Event& ev);

}  // namespace sdlgen
