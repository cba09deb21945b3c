// clb_geometry.cpp -- seeded synthetic media for the cases (complab3d_geometry.hh).
//
// The random stream is complab_geom::Rng (complab3d_geometry.hh:77-86, an
// xorshift64* generator) so a seed names the same packing here and in the
// reference's own generator.  clb_geom_pack adds the CompLaB material
// conventions the reference's ABC geometry tool writes
// (ABC_sweep_package/tools/make_geometries.py): a bounce-back shell on every
// grain, bounce-back y/z walls, isolated pockets walled off, percolation along
// x required, and the duplicated ghost x columns of readGeometry
// (complab_functions.hh:178-207).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "clb.h"

namespace {

class XorShift {
public:
    explicit XorShift(uint64_t seed) : s_(seed ? seed : 1) {}
    uint64_t next() {
        s_ ^= s_ >> 12;
        s_ ^= s_ << 25;
        s_ ^= s_ >> 27;
        return s_ * 2685821657736338717ULL;
    }
    double uniform() { return (double) (next() >> 11) * (1.0 / 9007199254740992.0); }
    int below(int n) { return n > 0 ? (int) (uniform() * (double) n) : 0; }

private:
    uint64_t s_;
};

struct Box {
    int nx, ny, nz;
    std::vector<int8_t> v;
    Box(int x, int y, int z, int8_t fill) : nx(x), ny(y), nz(z), v((size_t) x * y * z, fill) {}
    size_t at(int x, int y, int z) const { return (size_t) x + (size_t) nx * ((size_t) y + (size_t) ny * z); }
    bool in(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz; }
};

const int FACE[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

// carves one ball (or an x-y disc for cylinders) of solid; returns pore voxels removed
long carve(Box &b, int cx, int cy, int cz, double r, bool cyl, bool boundedLoops) {
    long removed = 0;
    const int R = (int) std::ceil(r);
    const int z0 = cyl ? 0 : (boundedLoops ? std::max(0, cz - R) : 0);
    const int z1 = cyl ? b.nz - 1 : (boundedLoops ? std::min(b.nz - 1, cz + R) : b.nz - 1);
    const int y0 = boundedLoops ? std::max(0, cy - R) : 0, y1 = boundedLoops ? std::min(b.ny - 1, cy + R) : b.ny - 1;
    const int x0 = boundedLoops ? std::max(0, cx - R) : 0, x1 = boundedLoops ? std::min(b.nx - 1, cx + R) : b.nx - 1;
    for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
                const double dx = x - cx, dy = y - cy, dz = cyl ? 0.0 : (double) (z - cz);
                if (dx * dx + dy * dy + dz * dz <= r * r) {
                    int8_t &m = b.v[b.at(x, y, z)];
                    if (m == CLB_PORE) ++removed;
                    if (m != CLB_WALL || !boundedLoops) m = CLB_SOLID;
                }
            }
    return removed;
}

// 27-point periodic box mean, `passes` times (the smoothing of the random-field kind)
void smoothField(std::vector<double> &f, int nx, int ny, int nz, int passes) {
    std::vector<double> t(f.size());
    for (int p = 0; p < passes; ++p) {
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    double s = 0.0;
                    int n = 0;
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                const int X = (x + dx + nx) % nx, Y = (y + dy + ny) % ny, Z = (z + dz + nz) % nz;
                                s += f[(size_t) X + (size_t) nx * ((size_t) Y + (size_t) ny * Z)];
                                ++n;
                            }
                    t[(size_t) x + (size_t) nx * ((size_t) y + (size_t) ny * z)] = s / (double) n;
                }
        f.swap(t);
    }
}

// the seeded media of complab_geom::generate (complab3d_geometry.hh:138-202)
bool generateInto(Box &b, int kind, double radius, int count, double porosity, double aperture, double roughness,
                  uint64_t seed) {
    const int nx = b.nx, ny = b.ny, nz = b.nz;
    XorShift rng(seed);
    switch (kind) {
    case CLB_GEOM_CHANNEL:
        return true;
    case CLB_GEOM_SPHERES:
    case CLB_GEOM_CYLINDERS: {
        const bool cyl = kind == CLB_GEOM_CYLINDERS;
        for (int k = 0; k < count; ++k) {
            const int cx = rng.below(nx), cy = rng.below(ny);
            const int cz = cyl ? 0 : rng.below(nz);
            carve(b, cx, cy, cz, radius, cyl, false);
        }
        return true;
    }
    case CLB_GEOM_RANDOM: {
        // a smoothed uniform field thresholded at the quantile of the porosity
        std::vector<double> f(b.v.size());
        for (double &x : f) x = rng.uniform();
        smoothField(f, nx, ny, nz, 2);
        std::vector<double> sorted(f);
        std::sort(sorted.begin(), sorted.end());
        const size_t cut = (size_t) ((1.0 - porosity) * (double) sorted.size());
        const double thr = sorted[cut < sorted.size() ? cut : sorted.size() - 1];
        for (size_t i = 0; i < f.size(); ++i) b.v[i] = f[i] >= thr ? CLB_PORE : CLB_SOLID;
        return true;
    }
    case CLB_GEOM_FRACTURE: {
        // a rough aperture around y = ny/2, three sine modes per z slice
        const double twoPi = 6.283185307179586;
        for (int8_t &m : b.v) m = CLB_SOLID;
        const double mid = 0.5 * ny;
        std::vector<double> wob((size_t) nx);
        for (int z = 0; z < nz; ++z) {
            std::fill(wob.begin(), wob.end(), 0.0);
            for (int k = 1; k <= 3; ++k) {
                const double ph = rng.uniform() * twoPi;
                for (int x = 0; x < nx; ++x) wob[(size_t) x] += (roughness / k) * std::sin(twoPi * k * x / nx + ph);
            }
            for (int x = 0; x < nx; ++x)
                for (int y = 0; y < ny; ++y)
                    if (std::fabs(y - (mid + wob[(size_t) x])) <= 0.5 * aperture) b.v[b.at(x, y, z)] = CLB_PORE;
        }
        return true;
    }
    }
    return false;
}

// addWalls (complab3d_geometry.hh:92-103): bit 1 x, 2 y, 4 z
void addWalls(Box &b, int axes) {
    for (int z = 0; z < b.nz; ++z)
        for (int y = 0; y < b.ny; ++y)
            for (int x = 0; x < b.nx; ++x) {
                const bool onX = (x == 0 || x == b.nx - 1), onY = (y == 0 || y == b.ny - 1), onZ = (z == 0 || z == b.nz - 1);
                if (((axes & 1) && onX) || ((axes & 2) && onY) || ((axes & 4) && onZ)) b.v[b.at(x, y, z)] = CLB_WALL;
            }
}

}  // namespace

extern "C" {

int clb_geom_generate(int nx, int ny, int nz, int kind, double radius, int count, double porosity, double aperture,
                      double roughness, uint64_t seed, int walls, int8_t *out) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || count < 0 || !out) return CLB_E_ARG;
    Box b(nx, ny, nz, CLB_PORE);
    if (!generateInto(b, kind, radius, count, porosity, aperture, roughness, seed)) return CLB_E_ARG;
    addWalls(b, walls);
    for (size_t i = 0; i < b.v.size(); ++i) out[i] = b.v[i];
    return CLB_OK;
}

int clb_rng_draws(uint64_t seed, int n, uint64_t *out) {
    if (n < 0 || (n > 0 && !out)) return CLB_E_ARG;
    XorShift r(seed);
    for (int k = 0; k < n; ++k) out[k] = r.next();
    return CLB_OK;
}

int clb_geom_spheres(int nx, int ny, int nz, double radius, int count, uint64_t seed, int walls_yz, int8_t *out) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || count < 0 || !out) return CLB_E_ARG;
    Box b(nx, ny, nz, CLB_PORE);
    XorShift rng(seed);
    for (int k = 0; k < count; ++k) {
        const int cx = rng.below(nx), cy = rng.below(ny), cz = rng.below(nz);
        carve(b, cx, cy, cz, radius, false, false);
    }
    if (walls_yz)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x)
                    if (y == 0 || y == ny - 1 || z == 0 || z == nz - 1) b.v[b.at(x, y, z)] = CLB_WALL;
    for (size_t i = 0; i < b.v.size(); ++i) out[i] = b.v[i];
    return CLB_OK;
}

int clb_geom_pack(int nxf, int ny, int nz, double radius, double porosity, uint64_t seed, int kind,
                  int8_t *lat, double *porosity_out) {
    if (nxf <= 1 || ny <= 2 || nz <= 2 || radius <= 0 || !lat) return CLB_E_ARG;
    if (kind != CLB_GEOM_SPHERES && kind != CLB_GEOM_CYLINDERS && kind != CLB_GEOM_RANDOM && kind != CLB_GEOM_FRACTURE)
        return CLB_E_ARG;
    Box b(nxf, ny, nz, CLB_PORE);
    const double total = (double) b.v.size();
    if (kind == CLB_GEOM_RANDOM) {
        generateInto(b, kind, radius, 0, porosity, 0.0, 0.0, seed);
        addWalls(b, 6);
    } else if (kind == CLB_GEOM_FRACTURE) {
        // aperture from the porosity of the slab (walls excluded), roughness radius / 2
        generateInto(b, kind, radius, 0, porosity, std::max(2.0, porosity * (ny - 2)), 0.5 * radius, seed);
        addWalls(b, 6);
    } else {
        addWalls(b, 6);
        long open = 0;
        for (int8_t m : b.v) open += (m == CLB_PORE);
        XorShift rng(seed);
        const bool cyl = (kind == CLB_GEOM_CYLINDERS);
        for (int guard = 0; (double) open / total > porosity && guard < 10000000; ++guard) {
            const int cx = rng.below(nxf), cy = rng.below(ny), cz = rng.below(nz);
            open -= carve(b, cx, cy, cz, radius, cyl, true);
        }
    }
    // flood fill: keep the pore cluster(s) touching both x faces
    std::vector<int> label(b.v.size(), 0);
    std::vector<char> spans(1, 0);
    std::vector<size_t> stack;
    int nlab = 0;
    for (size_t s0 = 0; s0 < b.v.size(); ++s0) {
        if (b.v[s0] != CLB_PORE || label[s0]) continue;
        ++nlab;
        bool west = false, east = false;
        stack.assign(1, s0);
        label[s0] = nlab;
        while (!stack.empty()) {
            const size_t i = stack.back();
            stack.pop_back();
            const int x = (int) (i % (size_t) nxf), y = (int) ((i / (size_t) nxf) % (size_t) ny), z = (int) (i / ((size_t) nxf * ny));
            west |= (x == 0);
            east |= (x == nxf - 1);
            for (const auto &d : FACE) {
                const int X = x + d[0], Y = y + d[1], Z = z + d[2];
                if (!b.in(X, Y, Z)) continue;
                const size_t j = b.at(X, Y, Z);
                if (b.v[j] == CLB_PORE && !label[j]) { label[j] = nlab; stack.push_back(j); }
            }
        }
        spans.push_back(west && east);
    }
    bool percolates = false;
    for (size_t i = 0; i < b.v.size(); ++i) {
        if (b.v[i] != CLB_PORE) continue;
        if (spans[label[i]]) percolates = true;
        else b.v[i] = CLB_WALL;
    }
    // a solid voxel that touches pore through a face becomes the grain's bounce-back shell
    std::vector<int8_t> shell = b.v;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nxf; ++x) {
                if (b.v[b.at(x, y, z)] != CLB_SOLID) continue;
                for (const auto &d : FACE) {
                    const int X = x + d[0], Y = y + d[1], Z = z + d[2];
                    if (b.in(X, Y, Z) && b.v[b.at(X, Y, Z)] == CLB_PORE) { shell[b.at(x, y, z)] = CLB_WALL; break; }
                }
            }
    long pore = 0;
    for (int8_t m : shell) pore += (m == CLB_PORE);
    if (porosity_out) *porosity_out = (double) pore / total;
    const int nx = nxf + 2;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int xs = x == 0 ? 0 : (x == nx - 1 ? nxf - 1 : x - 1);
                lat[(size_t) x + (size_t) nx * ((size_t) y + (size_t) ny * z)] = shell[b.at(xs, y, z)];
            }
    return percolates ? CLB_OK : CLB_E_GEOMETRY;
}

}  // extern "C"
