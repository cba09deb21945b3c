/*
 * ref_geom.cpp -- TEST INFRASTRUCTURE.  Builds the reference's geometry
 * generator (src/complab3d_geometry.hh, which needs only the C++ standard
 * library) from its own source, with nothing standing in for anything, and
 * exposes it through extern "C" so the tests can pin the oracle's generator
 * and the device's against it.  Built into oracle/_ref/libref_geom.so by
 * oracle/Makefile from the source where it lies; nothing of the reference is
 * copied into this repository.
 */
#include "complab3d_geometry.hh"

#include <cstdint>

extern "C" {

/* complab_geom::Rng draws */
int ref_rng_draws(uint64_t seed, int n, uint64_t *out) {
    complab_geom::Rng r(seed);
    for (int k = 0; k < n; ++k) out[k] = r.next();
    return 0;
}

/* complab_geom::generate; kind 0 spheres, 1 random, 2 cylinders, 3 fracture.
 * out: nx*ny*nz codes, index x + nx*(y + ny*z) */
int ref_geom_generate(int nx, int ny, int nz, int kind, double radius, int count, double porosity,
                      uint64_t seed, const char *walls, int8_t *out) {
    complab_geom::GenOptions o;
    const char *kinds[4] = {"spheres", "random", "cylinders", "fracture"};
    if (kind < 0 || kind > 3) return -1;
    o.kind = kinds[kind];
    o.radius = radius; o.count = count; o.porosity = porosity; o.seed = seed;
    o.walls = walls ? walls : "none";
    complab_geom::Grid g = complab_geom::generate(nx, ny, nz, o);
    if (g.size() == 0) return -1;
    for (size_t i = 0; i < g.size(); ++i) out[i] = (int8_t) g.v[i];
    return 0;
}

/* complab_geom::inspect: porosity, open voxels, clusters, spanning clusters, percolates */
int ref_geom_inspect(int nx, int ny, int nz, const int8_t *codes, double *porosity, long *open,
                     int *clusters, int *spanning, int *percolates) {
    complab_geom::Grid g(nx, ny, nz, 0);
    for (size_t i = 0; i < g.size(); ++i) g.v[i] = codes[i];
    complab_geom::Inspection R = complab_geom::inspect(g);
    *porosity = R.porosity; *open = R.openVoxels; *clusters = R.clusters;
    *spanning = R.spanningClusters; *percolates = R.percolates ? 1 : 0;
    return 0;
}

}  // extern "C"
