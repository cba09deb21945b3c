// clb_run_processors: coupled steps driven from C++ in the reference's
// processor form, through include/clb_processors.hh alone -- the calls the
// reference's main loop makes (complab.cpp:2307-2351, the CA block of
// 2793-3055), each process() forwarding to the device.
//
//   clb_run_processors <params.bin> <geometry.bin> <init.bin> <steps> <out.bin>
//
// params.bin is a clb_case_params as laid out in clb.h; geometry.bin the
// nx*ny*nz material codes (int8, x fastest); init.bin holds float64 arrays of
// nx*ny*nz values: the three velocity components, every species'
// concentration, every microbe's biomass, then omega_pore and omega_biofilm
// (n_species values each).  Per step: collide + stream of every transport
// lattice, run_kinetics + update_rxnLattices, run_abiotic_kinetics +
// update_abiotic_rxnLattices, then updateLocalMaskNtotalLattices3D and, while
// the largest biomass total exceeds Bmax, the spill rule of the case with
// updateAgeDistance3D relaxed to its fixed point every 50th sweep.  out.bin
// receives float64 fields: every species, every microbe's biomass, the mask,
// the age.  Prints the CA sweeps run.  Exit status 0, or 1 with the message.
#include "clb_processors.hh"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

using namespace clb_proc;

static int die(clb_ctx *ctx, const char *what, int rc) {
    std::fprintf(stderr, "%s: status %d: %s\n", what, rc, ctx ? clb_last_error(ctx) : "no context");
    return 1;
}

int main(int argc, char **argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s params.bin geometry.bin init.bin steps out.bin\n", argv[0]);
        return 1;
    }
    clb_case_params p;
    {
        std::ifstream f(argv[1], std::ios::binary);
        if (!f.read(reinterpret_cast<char *>(&p), sizeof(p))) { std::fprintf(stderr, "cannot read %s\n", argv[1]); return 1; }
    }
    const size_t N = (size_t) p.nx * p.ny * p.nz;
    const int S = p.n_species, M = p.n_microbes;
    std::vector<int8_t> geo(N);
    {
        std::ifstream f(argv[2], std::ios::binary);
        if (!f.read(reinterpret_cast<char *>(geo.data()), (std::streamsize) N)) { std::fprintf(stderr, "cannot read %s\n", argv[2]); return 1; }
    }
    std::vector<double> init((3 + S + M) * N + 2 * (size_t) S);
    {
        std::ifstream f(argv[3], std::ios::binary);
        if (!f.read(reinterpret_cast<char *>(init.data()), (std::streamsize) (init.size() * sizeof(double)))) {
            std::fprintf(stderr, "cannot read %s\n", argv[3]);
            return 1;
        }
    }
    const int steps = std::atoi(argv[4]);

    clb_ctx *ctx = nullptr;
    int rc = clb_ctx_create(0, &ctx);
    if (rc != CLB_OK) return die(ctx, "clb_ctx_create", rc);
    int id = -1;
    if ((rc = clb_case_create(ctx, &p, geo.data(), nullptr, &id)) != CLB_OK) return die(ctx, "clb_case_create", rc);
    const Target t = {ctx, id};
    const double *src = init.data();
    for (int k = 0; k < 3; ++k, src += N)
        if ((rc = clb_write_field(ctx, id, CLB_FIELD_VEL, k, src)) != CLB_OK) return die(ctx, "velocity", rc);
    for (int s = 0; s < S; ++s, src += N)
        if ((rc = clb_write_field(ctx, id, CLB_FIELD_CONC, s, src)) != CLB_OK) return die(ctx, "concentration", rc);
    for (int m = 0; m < M; ++m, src += N)
        if ((rc = clb_write_field(ctx, id, CLB_FIELD_BIOMASS, m, src)) != CLB_OK) return die(ctx, "biomass", rc);
    if ((rc = clb_set_ade_relaxation(ctx, id, src, src + S, p.dt)) != CLB_OK) return die(ctx, "relaxation", rc);

    // the constructor arguments the reference passes (complab.cpp:2339-2351, 2793-2800)
    std::vector<std::vector<double> > Kc((size_t) M, std::vector<double>((size_t) S, 0.0));
    std::vector<double> mu((size_t) M);
    for (int m = 0; m < M; ++m) {
        mu[(size_t) m] = p.mu_max[m];
        if (p.donor >= 0) Kc[(size_t) m][(size_t) p.donor] = p.ks[m];
        if (p.acceptor >= 0) Kc[(size_t) m][(size_t) p.acceptor] = p.ke[m];
    }
    std::vector<std::vector<plint> > bio((size_t) M, std::vector<plint>(1));
    for (int m = 0; m < M; ++m) bio[(size_t) m][0] = CLB_BIOFILM + m;
    const std::vector<plint> pore(1, CLB_PORE);
    const plint nx = p.nx, ny = p.ny, nz = p.nz, length = p.nx - 2;
    updateLocalMaskNtotalLattices3D<double> mt(nx, ny, nz, length, CLB_WALL, CLB_SOLID, bio, pore, p.thrd_bfilm_frac, p.bmax);
    updateAgeDistance3D<double> age(p.bmax, nx, ny, nz);

    long sweeps = 0;
    for (int it = 0; it < steps; ++it) {
        if ((rc = clb_ade_collide_stream(ctx, id, 1)) != CLB_OK) return die(ctx, "collide and stream", rc);
        if (M > 0 && p.bio_law != CLB_BIO_NONE) {
            if ((rc = applyProcessingFunctional(new run_kinetics<double>(nx, S, M, p.dt, Kc, mu, CLB_SOLID, CLB_WALL), t)))
                return die(ctx, "run_kinetics", rc);
            if ((rc = applyProcessingFunctional(new update_rxnLattices<double>(nx, S, M, CLB_SOLID, CLB_WALL), t)))
                return die(ctx, "update_rxnLattices", rc);
        }
        if (p.n_rxn > 0) {
            if ((rc = applyProcessingFunctional(new run_abiotic_kinetics<double>(nx, S, p.dt, CLB_SOLID, CLB_WALL), t)))
                return die(ctx, "run_abiotic_kinetics", rc);
            if ((rc = applyProcessingFunctional(new update_abiotic_rxnLattices<double>(nx, S, CLB_SOLID, CLB_WALL), t)))
                return die(ctx, "update_abiotic_rxnLattices", rc);
        }
        if (M <= 0) continue;
        if ((rc = applyProcessingFunctional(mt, t))) return die(ctx, "updateLocalMaskNtotalLattices3D", rc);
        double prev = mt.getMaxTotal();
        int stall = 0;
        for (long wc = 0; mt.getMaxTotal() - p.bmax > 1e-12; ++wc) {
            if (p.ca_method == CLB_CA_HALF)
                rc = applyProcessingFunctional(new halfPushExcessBiomass3D<double>(p.bmax, nx, ny, nz, 1, length, CLB_SOLID, CLB_WALL, pore), t);
            else
                rc = applyProcessingFunctional(new pushExcessBiomass3D<double>(p.bmax, nx, ny, nz, 1, length, CLB_SOLID, CLB_WALL, pore), t);
            if (rc) return die(ctx, "pushExcessBiomass3D", rc);
            if ((rc = applyProcessingFunctional(mt, t))) return die(ctx, "updateLocalMaskNtotalLattices3D", rc);
            ++sweeps;
            if (mt.getMaxTotal() > prev - 1e-12) { if (++stall >= 100) return die(ctx, "CA spill stalled", CLB_E_CA_STALL); }
            else stall = 0;
            prev = mt.getMaxTotal();
            if (wc % 50 == 0 && (rc = age.relaxToFixedPoint(t))) return die(ctx, "updateAgeDistance3D", rc);
            if (wc > 1000) return die(ctx, "CA spill did not settle", CLB_E_CA_SWEEPS);
        }
    }
    std::printf("sweeps %ld\n", sweeps);

    std::ofstream out(argv[5], std::ios::binary);
    std::vector<double> field(N);
    for (int s = 0; s < S; ++s) {
        if ((rc = clb_read_field(ctx, id, CLB_FIELD_CONC, s, field.data()))) return die(ctx, "read", rc);
        out.write(reinterpret_cast<const char *>(field.data()), (std::streamsize) (N * sizeof(double)));
    }
    for (int m = 0; m < M; ++m) {
        if ((rc = clb_read_field(ctx, id, CLB_FIELD_BIOMASS, m, field.data()))) return die(ctx, "read", rc);
        out.write(reinterpret_cast<const char *>(field.data()), (std::streamsize) (N * sizeof(double)));
    }
    const int dense[2] = {CLB_FIELD_MASK, CLB_FIELD_AGE};
    for (int f : dense) {
        if ((rc = clb_read_field(ctx, id, f, 0, field.data()))) return die(ctx, "read", rc);
        out.write(reinterpret_cast<const char *>(field.data()), (std::streamsize) (N * sizeof(double)));
    }
    clb_case_destroy(ctx, id);
    clb_ctx_destroy(ctx);
    return 0;
}
