// Synthetic detector, connection-map builder and seeded event generator.
// See sdl_gen.h for what is generated and why.
#include "sdl_gen.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>

namespace sdlgen {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSlopeInf = 123456789.0;           // "infinite" slope marker of the geometry files
constexpr double kCmPerGeV = 100.0 / (0.299792458 * 3.8);  // helix radius [cm] per GeV of pT at 3.8 T

// ---------------------------------------------------------------- detId helpers
// Field layout of Module.cc:159-235 / Module.h:31-46.
uint32_t barrelBase(int layer, int side, int rod, int module) {
    return (5u << 25) | (uint32_t(layer) << 20) | (uint32_t(side) << 18) | (uint32_t(rod) << 10) |
           (uint32_t(module) << 2);
}
uint32_t endcapBase(int side, int layer, int ring, int module) {
    return (4u << 25) | (uint32_t(side) << 23) | (uint32_t(layer) << 18) | (uint32_t(ring) << 12) |
           (uint32_t(module) << 2);
}
// Inversion rule of Module.cc:239-302.
bool invertedRule(bool barrel, int side, int layer, int module) {
    if (!barrel) return side == 1 ? (module % 2 == 1) : (module % 2 == 0);
    if (side == 3) return layer <= 3 ? (module % 2 == 1) : (module % 2 == 0);
    return layer <= 2 ? (module % 2 == 1) : (module % 2 == 0);
}
// Low bits chosen so that the partner is detId +/- 1 without a carry into the
// module field: not inverted -> lower ..01 / upper ..10; inverted -> lower ..10 / upper ..01.
void lowerUpper(uint32_t base, bool inverted, uint32_t& lower, uint32_t& upper) {
    if (!inverted) { lower = base | 1u; upper = base | 2u; }
    else           { lower = base | 2u; upper = base | 1u; }
}
// Sensor separations assumed by the cut algebra (MiniDoublet.cc:1260-1332).
bool tighterTilted(int side, int layer, int rod) {
    if (layer == 3) return true;
    if (side == 1 && layer == 2 && rod > 5) return true;
    if (side == 2 && layer == 2 && rod < 8) return true;
    if (side == 1 && layer == 1 && rod > 9) return true;
    if (side == 2 && layer == 1 && rod < 4) return true;
    return false;
}
double endcapGap(int layer, int ring) {
    int j = ring - 1;
    if (layer <= 2) return j < 10 ? 0.4 : 0.18;
    if (layer <= 4) return j < 8 ? 0.4 : 0.18;
    return j < 9 ? 0.4 : 0.18;
}

// --------------------------------------------------------------------- RNG
inline uint64_t splitmix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
        s = splitmix(seed);
        s = splitmix(s ^ (a * 0x100000001b3ULL));
        s = splitmix(s ^ (b * 0x9e3779b97f4a7c15ULL));
        s = splitmix(s ^ (c + 0x632be59bd9b4e019ULL));
    }
    uint64_t next() { s += 0x9e3779b97f4a7c15ULL; return splitmix(s); }
    double uni() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uni(double a, double b) { return a + (b - a) * uni(); }
    double gauss() {
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
    int poisson(double lam) {
        double L = std::exp(-lam), p = 1.0;
        int k = 0;
        do { ++k; p *= uni(); } while (p > L);
        return k - 1;
    }
};

// ------------------------------------------------------------- geometry build
double slopeOf(int k, int n) {
    // x-y slope of the tangential direction (-sin phi, cos phi) at phi = 2 pi k / n.
    if (2 * k == n || k == 0) return kSlopeInf;
    if (4 * k == n || 4 * k == 3 * n) return 0.0;
    double phi = 2.0 * kPi * k / n;
    return -std::cos(phi) / std::sin(phi);
}

struct Builder {
    Geometry& g;
    explicit Builder(Geometry& gg) : g(gg) {}

    int addModule(uint32_t base, bool inverted, const double c[3], const double n[3], const double u[3],
                  const double v[3], double hu, double hv, double gap, bool ps, bool barrel, bool tilted,
                  int layer, int surf, double drdz, double slope, double cphi, double cr, double cz) {
        uint32_t lo, up;
        lowerUpper(base, inverted, lo, up);
        Module m{};
        m.surf = surf; m.barrel = barrel; m.ps = ps; m.tilted = tilted; m.layer = layer;
        m.drdz = drdz; m.slope = slope; m.centroidPhi = cphi; m.centroidR = cr; m.centroidZ = cz;
        int mi = int(g.modules.size());
        for (int s = 0; s < 2; ++s) {
            Sensor se{};
            se.detId = s == 0 ? lo : up;
            double off = (s == 0 ? -0.5 : 0.5) * gap;
            for (int i = 0; i < 3; ++i) {
                se.c[i] = c[i] + off * n[i];
                se.n[i] = n[i]; se.u[i] = u[i]; se.v[i] = v[i];
            }
            se.hu = hu; se.hv = hv;
            // Pixel / strip assignment of Module.cc:340-358: the lower sensor is the
            // pixel one unless the module is inverted.
            bool pixel = ps && ((s == 0) != inverted);
            const double r12 = 1.0 / std::sqrt(12.0);
            if (pixel)      { se.sig_u = 0.0100 * r12; se.sig_v = 0.15 * r12; }
            else if (ps)    { se.sig_u = 0.0090 * r12; se.sig_v = 2.5 * r12; }
            else            { se.sig_u = 0.0090 * r12; se.sig_v = 5.0 * r12; }
            se.module = mi;
            m.sensor[s] = int(g.sensors.size());
            g.sensors.push_back(se);
        }
        g.modules.push_back(m);
        return mi;
    }

    void barrelLayer(int L, double R, bool full) {
        Surface S;
        S.barrel = true; S.rz = R; S.layer = L;
        int si = int(g.surfs.size());
        const int nphi = int(std::ceil(2.0 * kPi * R / 10.0));
        const bool ps = L <= 3;
        const double flatGap = 0.26;
        // flat section
        double zmax = (full && L <= 3) ? 15.0 : 120.0;
        int nz = int(std::lround(2.0 * zmax / 5.0));
        for (int j = 0; j < nz; ++j) {
            Ring ring;
            ring.pos = -zmax + 2.5 + 5.0 * j; ring.half = 2.55; ring.nphi = nphi; ring.phi0 = 0.0;
            for (int k = 0; k < nphi; ++k) {
                double phi = 2.0 * kPi * k / nphi, cp = std::cos(phi), sp = std::sin(phi);
                double c[3] = {R * cp, R * sp, ring.pos}, n[3] = {cp, sp, 0}, u[3] = {-sp, cp, 0}, v[3] = {0, 0, 1};
                int rod = k + 1, mod = j + 1;
                bool inv = invertedRule(true, 3, L, mod);
                int mi = addModule(barrelBase(L, 3, rod, mod), inv, c, n, u, v, 5.05, 2.55, flatGap, ps, true,
                                   false, L, si, 0.0, slopeOf(k, nphi), phi, R, ring.pos);
                ring.modules.push_back(mi);
            }
            S.rings.push_back(ring);
        }
        if (full && L <= 3) {
            // Tilted rings, 15 < |z| < 120, tilt B = min(atan(z/R), 60 deg).  Rings are
            // placed projectively: the near end of each ring sits on the line from the
            // interaction point through the far end of the previous ring (with a small
            // angular overlap), so a track from the origin crosses about one ring.
            std::vector<std::pair<double, double>> zr;  // (zc, tilt)
            double alpha = std::atan2(R, 15.0);          // polar angle already covered
            const double ov = 0.004;                     // rad of overlap between rings
            for (int guard = 0; guard < 120; ++guard) {
                double zest = R / std::tan(alpha) + 2.0;
                double B = std::min(std::atan(zest / R), 60.0 * kPi / 180.0);
                double cB = std::cos(B), sB = std::sin(B);
                double zc = (R + 2.5 * sB) / std::tan(alpha + ov) + 2.5 * cB;
                if (zc - 2.5 * cB > 120.0) break;
                zr.push_back({zc, B});
                alpha = std::atan2(R - 2.5 * sB, zc + 2.5 * cB);
            }
            int nr = int(zr.size());
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                int side = sgn > 0 ? 2 : 1;
                for (int k = 0; k < nr; ++k) {
                    double zc = sgn * zr[k].first, B = zr[k].second;
                    int rod = sgn > 0 ? k + 1 : nr - k;  // PosZ counts outward, NegZ counts toward z=0
                    Ring ring;
                    ring.pos = zc; ring.half = 2.5 * std::cos(B) + 0.1; ring.nphi = nphi; ring.phi0 = 0.0;
                    double gap = tighterTilted(side, L, rod) ? 0.26 : 0.4;
                    for (int p = 0; p < nphi; ++p) {
                        double phi = 2.0 * kPi * p / nphi, cp = std::cos(phi), sp = std::sin(phi);
                        double cB = std::cos(B), sB = std::sin(B);
                        double c[3] = {R * cp, R * sp, zc};
                        double n[3] = {cp * cB, sp * cB, sgn * sB};
                        double u[3] = {-sp, cp, 0};
                        double v[3] = {-sgn * sB * cp, -sgn * sB * sp, cB};
                        int mod = p + 1;
                        bool inv = invertedRule(true, side, L, mod);
                        int mi = addModule(barrelBase(L, side, rod, mod), inv, c, n, u, v, 5.05, 2.55, gap, ps,
                                           true, true, L, si, std::tan(B), slopeOf(p, nphi), phi, R, zc);
                        ring.modules.push_back(mi);
                    }
                    S.rings.push_back(ring);
                }
            }
        }
        g.surfs.push_back(S);
    }

    void endcapDisk(int d, double Zabs, int sgn) {
        Surface S;
        S.barrel = false; S.rz = sgn * Zabs; S.layer = d;
        int si = int(g.surfs.size());
        int side = sgn > 0 ? 2 : 1;
        int nps = d <= 2 ? 10 : 7;
        double rmin = std::max(22.0, Zabs / std::sinh(2.5));
        std::vector<std::pair<double, double>> rings;  // centre, half
        double w = (85.0 - rmin) / nps;
        for (int k = 0; k < nps; ++k) rings.push_back({rmin + w * (k + 0.5), 0.5 * w + 0.15});
        for (int k = 0; k < 5; ++k) rings.push_back({87.5 + 4.9 * k, 2.5});
        for (int r = 0; r < int(rings.size()); ++r) {
            int ringNo = r + 1;
            bool ps = ringNo <= (d <= 2 ? 10 : 7);
            double rc = rings[r].first, half = rings[r].second;
            int nphi = int(std::ceil(2.0 * kPi * (rc + half) / 10.0));
            Ring ring;
            ring.pos = rc; ring.half = half; ring.nphi = nphi; ring.phi0 = 0.0;
            double gap = endcapGap(d, ringNo);
            for (int p = 0; p < nphi; ++p) {
                double phi = 2.0 * kPi * p / nphi, cp = std::cos(phi), sp = std::sin(phi);
                double c[3] = {rc * cp, rc * sp, sgn * Zabs};
                double n[3] = {0, 0, double(sgn)}, u[3] = {-sp, cp, 0}, v[3] = {cp, sp, 0};
                int mod = p + 1;
                bool inv = invertedRule(false, side, d, mod);
                int mi = addModule(endcapBase(side, d, ringNo, mod), inv, c, n, u, v, 5.05, half, gap, ps, false,
                                   false, d, si, 0.0, slopeOf(p, nphi), phi, rc, sgn * Zabs);
                ring.modules.push_back(mi);
            }
            S.rings.push_back(ring);
        }
        g.surfs.push_back(S);
    }
};

// --------------------------------------------------------------- propagation
struct Helix {
    double x0, y0, z0, psi0, lam, R, h;  // h = -charge (turning sense), R in cm
    void at(double t, double p[3]) const {
        double psi = psi0 + h * t / R;
        p[0] = x0 + (R / h) * (std::sin(psi) - std::sin(psi0));
        p[1] = y0 - (R / h) * (std::cos(psi) - std::cos(psi0));
        p[2] = z0 + lam * t;
    }
    void dir(double t, double d[3]) const {
        double psi = psi0 + h * t / R;
        d[0] = std::cos(psi); d[1] = std::sin(psi); d[2] = lam;
    }
    // smallest transverse path t > tmin reaching radius rc
    double cylinder(double rc, double tmin) const {
        double cx = x0 - (R / h) * std::sin(psi0), cy = y0 + (R / h) * std::cos(psi0);
        double d = std::hypot(cx, cy);
        if (d > R + rc || d < std::fabs(R - rc) || d < 1e-12) return -1;
        double a = (rc * rc - R * R + d * d) / (2 * d);
        double hh2 = rc * rc - a * a;
        double hh = hh2 > 0 ? std::sqrt(hh2) : 0.0;
        double ex = cx / d, ey = cy / d;
        double best = -1;
        for (int s = -1; s <= 1; s += 2) {
            double px = a * ex - s * hh * ey, py = a * ey + s * hh * ex;
            double sn = h * (px - cx) / R, cs = -h * (py - cy) / R;
            double psi = std::atan2(sn, cs);
            double delta = h * (psi - psi0);
            delta = std::fmod(delta, 2 * kPi);
            if (delta < 0) delta += 2 * kPi;
            double t = delta * R;
            if (t <= tmin) t += 2 * kPi * R;
            if (best < 0 || t < best) best = t;
        }
        return best;
    }
    double disk(double z, double tmin) const {
        if (std::fabs(lam) < 1e-12) return -1;
        double t = (z - z0) / lam;
        return t > tmin ? t : -1;
    }
    // Newton on n.x(t) = n.c
    bool plane(const Sensor& s, double tguess, double& t) const {
        double dd = s.n[0] * s.c[0] + s.n[1] * s.c[1] + s.n[2] * s.c[2];
        t = tguess;
        for (int it = 0; it < 30; ++it) {
            double p[3], d[3];
            at(t, p); dir(t, d);
            double f = s.n[0] * p[0] + s.n[1] * p[1] + s.n[2] * p[2] - dd;
            double fp = s.n[0] * d[0] + s.n[1] * d[1] + s.n[2] * d[2];
            if (std::fabs(fp) < 1e-9) return false;
            double step = f / fp;
            if (step > 20) step = 20;
            if (step < -20) step = -20;
            t -= step;
            if (std::fabs(step) < 1e-10) return true;
        }
        return false;
    }
};

struct Crossing {
    int sensor;
    double p[3];
};

// Propagates a helix through the surfaces; calls visit(surfaceIndex, crossings, outward)
// for each layer crossing, in path order.  Returns after maxCross crossings.
template <class Visit, class Kick>
void propagate(const Geometry& g, Helix hx, Visit visit, Kick kick, int maxCross = 40) {
    double turned = 0.0;  // accumulated turning angle [rad]
    for (int nc = 0; nc < maxCross; ++nc) {
        double best = -1;
        int bestS = -1;
        for (int si = 0; si < int(g.surfs.size()); ++si) {
            const Surface& S = g.surfs[si];
            double t = S.barrel ? hx.cylinder(S.rz, 1e-3) : hx.disk(S.rz, 1e-3);
            if (t <= 0) continue;
            double p[3];
            hx.at(t, p);
            if (S.barrel) {
                if (std::fabs(p[2]) > 125.0) continue;
            } else {
                double r = std::hypot(p[0], p[1]);
                if (r > 115.0 || r < 15.0) continue;
            }
            if (best < 0 || t < best) { best = t; bestS = si; }
        }
        if (bestS < 0) return;
        const Surface& S = g.surfs[bestS];
        double p[3], d[3];
        hx.at(best, p);
        hx.dir(best, d);
        bool outward = S.barrel ? (p[0] * d[0] + p[1] * d[1] > 0) : (p[2] * d[2] > 0);
        double phi = std::atan2(p[1], p[0]);
        double key = S.barrel ? p[2] : std::hypot(p[0], p[1]);
        double margin = S.barrel ? 3.0 * std::fabs(hx.lam) + 2.0 : 3.0;
        std::vector<Crossing> hits;
        for (const Ring& ring : S.rings) {
            if (std::fabs(key - ring.pos) > ring.half + margin) continue;
            double dphi = 2 * kPi / ring.nphi;
            double f = (phi - ring.phi0) / dphi;
            int k0 = int(std::floor(f + 0.5));
            for (int dk = -1; dk <= 1; ++dk) {
                int k = ((k0 + dk) % ring.nphi + ring.nphi) % ring.nphi;
                const Module& m = g.modules[ring.modules[k]];
                for (int s = 0; s < 2; ++s) {
                    const Sensor& se = g.sensors[m.sensor[s]];
                    double t;
                    if (!hx.plane(se, best, t)) continue;
                    if (t <= 1e-6 || std::fabs(t - best) > 60.0) continue;
                    double q[3];
                    hx.at(t, q);
                    double a = 0, b = 0;
                    for (int i = 0; i < 3; ++i) { a += (q[i] - se.c[i]) * se.u[i]; b += (q[i] - se.c[i]) * se.v[i]; }
                    if (std::fabs(a) > se.hu || std::fabs(b) > se.hv) continue;
                    hits.push_back({m.sensor[s], {q[0], q[1], q[2]}});
                }
            }
        }
        if (!visit(bestS, hits, outward)) return;
        // The track leaves the tracker once it crosses the outermost barrel layer or
        // the outermost disk going outward; loopers stop after three quarters of a turn.
        if (outward && (S.barrel ? S.rz >= 109.0 : std::fabs(S.rz) >= 260.0)) return;
        turned += best / hx.R;
        if (turned > 1.5 * kPi) return;
        // restart the helix at the crossing point, optionally with a scattering kick
        double nd[3] = {d[0], d[1], d[2]};
        kick(nd);
        double tr = std::hypot(nd[0], nd[1]);
        if (tr < 1e-9) return;
        Helix nh = hx;
        nh.x0 = p[0]; nh.y0 = p[1]; nh.z0 = p[2];
        nh.psi0 = std::atan2(nd[1], nd[0]);
        double ptot_over_pt_old = std::sqrt(1 + hx.lam * hx.lam);
        double len = std::sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
        double pt_ratio = (tr / len) * ptot_over_pt_old;  // new pT / old pT at fixed |p|
        nh.R = hx.R * pt_ratio;
        nh.lam = nd[2] / tr;
        hx = nh;
    }
}

void buildConnectionMap(Geometry& g, int ntracks) {
    std::map<uint32_t, std::set<uint32_t>> conn;
    double etaMax = g.kind == GEOM_BARREL_FLAT ? 1.2 : 2.6;
    for (int it = 0; it < ntracks; ++it) {
        Rng rng(0xc0ffee1234ULL, 99, uint64_t(g.kind), uint64_t(it));
        double invpt = rng.uni(0.0, 1.0 / 0.8);
        double pt = 1.0 / std::max(invpt, 1e-3);
        double eta = rng.uni(-etaMax, etaMax);
        double phi = rng.uni(-kPi, kPi);
        double q = rng.uni() < 0.5 ? -1.0 : 1.0;
        Helix hx{0.0, 0.0, rng.uni(-15.0, 15.0), phi, std::sinh(eta), pt * kCmPerGeV, -q};
        std::vector<uint32_t> prev;
        propagate(
            g, hx,
            [&](int, const std::vector<Crossing>& cs, bool outward) {
                if (!outward) return false;
                std::vector<uint32_t> cur;
                for (const Crossing& c : cs) {
                    const Sensor& s = g.sensors[c.sensor];
                    uint32_t lower = g.sensors[g.modules[s.module].sensor[0]].detId;
                    if (std::find(cur.begin(), cur.end(), lower) == cur.end()) cur.push_back(lower);
                }
                if (cur.empty()) return true;
                for (uint32_t a : prev)
                    for (uint32_t b : cur) conn[a].insert(b);
                prev = cur;
                return true;
            },
            [](double*) {});
    }
    g.conn.clear();
    for (auto& kv : conn) g.conn.push_back({kv.first, std::vector<uint32_t>(kv.second.begin(), kv.second.end())});
}

}  // namespace

void buildGeometry(int kind, Geometry& g, int mapTracks) {
    g = Geometry();
    g.kind = kind;
    Builder b(g);
    const double radii[6] = {23, 36, 51, 68, 86, 110};
    for (int L = 1; L <= 6; ++L) b.barrelLayer(L, radii[L - 1], kind == GEOM_FULL);
    if (kind == GEOM_FULL) {
        const double zs[5] = {131, 155, 185, 219, 261};
        for (int sgn = -1; sgn <= 1; sgn += 2)
            for (int d = 1; d <= 5; ++d) b.endcapDisk(d, zs[d - 1], sgn);
    }
    buildConnectionMap(g, mapTracks);
}

void writeGeometryFiles(const Geometry& g, const std::string& tiltedPath, const std::string& endcapPath,
                        const std::string& mapPath) {
    FILE* ft = std::fopen(tiltedPath.c_str(), "w");
    FILE* fe = std::fopen(endcapPath.c_str(), "w");
    FILE* fm = std::fopen(mapPath.c_str(), "w");
    if (!ft || !fe || !fm) {
        if (ft) std::fclose(ft);
        if (fe) std::fclose(fe);
        if (fm) std::fclose(fm);
        return;
    }
    for (const Module& m : g.modules) {
        for (int s = 0; s < 2; ++s) {
            uint32_t id = g.sensors[m.sensor[s]].detId;
            if (m.barrel) {
                std::fprintf(ft, "%u %.9g %.9g\n", id, float(m.drdz), float(m.slope));
            } else {
                // columns: detid avgr2 yl sl yh sh centroid(phi) centroid(r) z  -- the reader
                // stores column 7 as the centroid phi (EndcapGeometry.cc:57-58).
                double r2 = m.centroidR * m.centroidR;
                double yint = m.slope == kSlopeInf ? 0.0 : -m.slope * m.centroidR * std::cos(m.centroidPhi) +
                                                                m.centroidR * std::sin(m.centroidPhi);
                std::fprintf(fe, "%u %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", id, float(r2), float(yint),
                             float(m.slope), float(yint), float(m.slope), float(m.centroidPhi),
                             float(m.centroidR), float(m.centroidZ));
            }
        }
    }
    for (const auto& kv : g.conn) {
        std::fprintf(fm, "%u %d", kv.first, int(kv.second.size()));
        for (uint32_t d : kv.second) std::fprintf(fm, " %u", d);
        std::fprintf(fm, "\n");
    }
    std::fclose(ft);
    std::fclose(fe);
    std::fclose(fm);
}

int geometryKindOf(int workload) {
    return (workload == 1 || workload == 4 || workload == 11) ? GEOM_BARREL_FLAT : GEOM_FULL;
}

EventConfig presetConfig(int workload) {
    EventConfig c;
    c.workload = workload;
    switch (workload) {
        case 1:
        case 4:
            break;  // defaults: 500 tracks, 1/pT ~ U[0.1,1], |eta| < 0.8, no MS, 5% noise
        case 2:
            c.ntracks_min = c.ntracks_max = 8000; c.n_hard = 20; c.pt_mode = 1; c.eta_max = 2.4; c.ms = true;
            break;
        case 3:
            c.ntracks_min = 2000; c.ntracks_max = 10000; c.n_hard = 20; c.pt_mode = 1; c.eta_max = 2.4; c.ms = true;
            break;
        case 5:
            c.ntracks_min = 2000; c.ntracks_max = 10000; c.n_hard = 20; c.pt_mode = 1; c.eta_max = 2.4; c.ms = true;
            c.frac_lowpt = 0.15; c.frac_displaced = 0.05; c.noise_rate = 0.10;
            break;
        // small slices with the same geometry and distributions, for parity tests
        // at sizes the CPU reference finishes in seconds
        case 11:  // flat barrel, 40 tracks: chains with accept-all (AllComb) stages
            c.ntracks_min = c.ntracks_max = 40;
            break;
        case 12:
            c.ntracks_min = c.ntracks_max = 300; c.n_hard = 2; c.pt_mode = 1; c.eta_max = 2.4; c.ms = true;
            break;
        case 15:
            c.ntracks_min = c.ntracks_max = 300; c.n_hard = 2; c.pt_mode = 1; c.eta_max = 2.4; c.ms = true;
            c.frac_lowpt = 0.15; c.frac_displaced = 0.05; c.noise_rate = 0.10;
            break;
        default:
            break;
    }
    return c;
}

void generateEvent(const Geometry& g, const EventConfig& cfg, uint64_t ievent, Event& ev) {
    ev = Event();
    Rng er(cfg.seed, uint64_t(cfg.workload), ievent, 0xffffffffULL);
    int ntr = cfg.ntracks_min;
    if (cfg.ntracks_max > cfg.ntracks_min) ntr += int(er.uni() * (cfg.ntracks_max - cfg.ntracks_min + 1));
    ev.ntracks = ntr + cfg.n_hard;
    auto addHit = [&](int sensor, const double p[3], int sim) {
        ev.x.push_back(float(p[0]));
        ev.y.push_back(float(p[1]));
        ev.z.push_back(float(p[2]));
        ev.detId.push_back(g.sensors[sensor].detId);
        ev.idx.push_back(int(ev.idx.size()));
        ev.simTrack.push_back(sim);
    };
    for (int it = 0; it < ev.ntracks; ++it) {
        Rng rng(cfg.seed, uint64_t(cfg.workload), ievent, uint64_t(it));
        double pt;
        bool hard = it >= ntr;
        bool lowpt = !hard && rng.uni() < cfg.frac_lowpt;
        bool displaced = !hard && rng.uni() < cfg.frac_displaced;
        if (hard) pt = rng.uni(5.0, 100.0);
        else if (lowpt) pt = rng.uni(cfg.lowpt_lo, cfg.lowpt_hi);
        else if (cfg.pt_mode == 0) pt = 1.0 / rng.uni(cfg.invpt_lo, cfg.invpt_hi);
        else {
            do { pt = -0.6 * std::log(std::max(rng.uni() * rng.uni(), 1e-300)); } while (pt <= 0.8);
        }
        double eta = rng.uni(-cfg.eta_max, cfg.eta_max);
        double phi = rng.uni(-kPi, kPi);
        double q = rng.uni() < 0.5 ? -1.0 : 1.0;
        double vz = cfg.sigma_z * rng.gauss();
        double vx = 0, vy = 0;
        if (displaced) {
            double rv = rng.uni(0.0, cfg.rv_max), pv = rng.uni(-kPi, kPi);
            vx = rv * std::cos(pv); vy = rv * std::sin(pv);
        }
        double ptot = pt * std::cosh(eta);
        double theta0 = 0.0136 / ptot * std::sqrt(0.02);
        Helix hx{vx, vy, vz, phi, std::sinh(eta), pt * kCmPerGeV, -q};
        propagate(
            g, hx,
            [&](int, const std::vector<Crossing>& cs, bool) {
                for (const Crossing& c : cs) {
                    const Sensor& s = g.sensors[c.sensor];
                    double du = s.sig_u * rng.gauss(), dv = s.sig_v * rng.gauss();
                    double p[3];
                    for (int i = 0; i < 3; ++i) p[i] = c.p[i] + du * s.u[i] + dv * s.v[i];
                    addHit(c.sensor, p, it);
                }
                return true;
            },
            [&](double* d) {
                if (!cfg.ms) return;
                double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                double t[3] = {d[0] / len, d[1] / len, d[2] / len};
                // two unit vectors orthogonal to t
                double a[3] = {-t[1], t[0], 0.0};
                double an = std::hypot(a[0], a[1]);
                if (an < 1e-9) { a[0] = 1; a[1] = 0; an = 1; }
                a[0] /= an; a[1] /= an;
                double b[3] = {t[1] * a[2] - t[2] * a[1], t[2] * a[0] - t[0] * a[2], t[0] * a[1] - t[1] * a[0]};
                double k1 = theta0 * rng.gauss(), k2 = theta0 * rng.gauss();
                for (int i = 0; i < 3; ++i) d[i] = t[i] + k1 * a[i] + k2 * b[i];
            });
    }
    // uniform noise hits, one independent stream per sensor
    if (cfg.noise_rate > 0) {
        for (int si = 0; si < int(g.sensors.size()); ++si) {
            Rng rng(cfg.seed, uint64_t(cfg.workload), ievent, 0x100000000ULL + uint64_t(si));
            int nn = rng.poisson(cfg.noise_rate);
            const Sensor& s = g.sensors[si];
            for (int k = 0; k < nn; ++k) {
                double a = rng.uni(-s.hu, s.hu), b = rng.uni(-s.hv, s.hv);
                double p[3];
                for (int i = 0; i < 3; ++i) p[i] = s.c[i] + a * s.u[i] + b * s.v[i];
                addHit(si, p, -1);
            }
        }
    }
}

}  // namespace sdlgen

// ------------------------------------------------------------------ C API
extern "C" {

void* sdlgen_geometry_new(int kind, int mapTracks) {
    auto* g = new sdlgen::Geometry();
    sdlgen::buildGeometry(kind, *g, mapTracks);
    return g;
}

void sdlgen_geometry_free(void* g) { delete static_cast<sdlgen::Geometry*>(g); }

int sdlgen_geometry_nsensors(void* g) { return int(static_cast<sdlgen::Geometry*>(g)->sensors.size()); }

int sdlgen_geometry_write(void* g, const char* tilted, const char* endcap, const char* map) {
    sdlgen::writeGeometryFiles(*static_cast<sdlgen::Geometry*>(g), tilted, endcap, map);
    return 0;
}

void* sdlgen_event_new(void* g, int workload, unsigned long long ievent, unsigned long long seed) {
    auto* ev = new sdlgen::Event();
    sdlgen::EventConfig cfg = sdlgen::presetConfig(workload);
    if (seed) cfg.seed = seed;
    sdlgen::generateEvent(*static_cast<sdlgen::Geometry*>(g), cfg, ievent, *ev);
    return ev;
}

int sdlgen_event_nhits(void* e) { return int(static_cast<sdlgen::Event*>(e)->x.size()); }

int sdlgen_event_ntracks(void* e) { return static_cast<sdlgen::Event*>(e)->ntracks; }
int sdlgen_geometry_kind(int workload) { return sdlgen::geometryKindOf(workload); }

void sdlgen_event_copy(void* e, float* x, float* y, float* z, unsigned int* detId, int* idx, int* sim) {
    auto* ev = static_cast<sdlgen::Event*>(e);
    size_t n = ev->x.size();
    if (n == 0) return;
    std::memcpy(x, ev->x.data(), n * sizeof(float));
    std::memcpy(y, ev->y.data(), n * sizeof(float));
    std::memcpy(z, ev->z.data(), n * sizeof(float));
    std::memcpy(detId, ev->detId.data(), n * sizeof(unsigned int));
    std::memcpy(idx, ev->idx.data(), n * sizeof(int));
    if (sim) std::memcpy(sim, ev->simTrack.data(), n * sizeof(int));
}

void sdlgen_event_free(void* e) { delete static_cast<sdlgen::Event*>(e); }

}  // extern "C"
