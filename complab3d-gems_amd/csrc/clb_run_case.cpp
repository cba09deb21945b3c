// clb_run_case: a coupled-step run driven from C++ alone -- case creation, the
// set-up of clb_driver.hh (flow solve, Peclet rescale, ADE relaxation,
// stabilization), the steps, and the species fields written out.
//
//   clb_run_case <params.bin> <geometry.bin> <settings.txt> <out.bin>
//
// params.bin is a clb_case_params as laid out in clb.h; geometry.bin the
// nx*ny*nz material codes (int8, x fastest); settings.txt holds "key value"
// lines: pe, char_length, delta_p, ns_tau, max_iter, rule, eps, stabilize, dx,
// steps, d_pore (one value per species, repeated key), d_biofilm (likewise).
// out.bin receives every species' concentration field (float64, species after
// species); the set-up report goes to stdout.  Exit status 0, or 1 with the
// library's message on stderr.
#include "clb_driver.hh"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

static int die(clb_ctx *ctx, const char *what) {
    std::fprintf(stderr, "%s: %s\n", what, ctx ? clb_last_error(ctx) : "no context");
    return 1;
}

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s params.bin geometry.bin settings.txt out.bin\n", argv[0]);
        return 1;
    }
    clb_case_params p;
    {
        std::ifstream f(argv[1], std::ios::binary);
        if (!f.read(reinterpret_cast<char *>(&p), sizeof(p))) { std::fprintf(stderr, "cannot read %s\n", argv[1]); return 1; }
    }
    const size_t N = (size_t) p.nx * p.ny * p.nz;
    std::vector<int8_t> geo(N);
    {
        std::ifstream f(argv[2], std::ios::binary);
        if (!f.read(reinterpret_cast<char *>(geo.data()), (std::streamsize) N)) { std::fprintf(stderr, "cannot read %s\n", argv[2]); return 1; }
    }
    clb_driver::FlowSetup fs;
    clb_driver::TransportSetup ts;
    int steps = 0;
    {
        std::ifstream f(argv[3]);
        std::string line;
        while (std::getline(f, line)) {
            std::istringstream ls(line);
            std::string k;
            double v;
            if (!(ls >> k >> v)) continue;
            if (k == "pe") fs.pe = v;
            else if (k == "char_length") fs.char_length = v;
            else if (k == "delta_p") fs.delta_p = v;
            else if (k == "ns_tau") fs.ns_tau = v;
            else if (k == "max_iter") fs.max_iter = (int) v;
            else if (k == "rule") fs.rule = (int) v;
            else if (k == "eps") fs.eps = v;
            else if (k == "has_biofilm") fs.has_biofilm = v != 0;
            else if (k == "stabilize") ts.stabilize_steps = (int) v;
            else if (k == "dx") ts.dx = v;
            else if (k == "steps") steps = (int) v;
            else if (k == "d_pore") ts.d_pore.push_back(v);
            else if (k == "d_biofilm") ts.d_biofilm.push_back(v);
        }
    }
    clb_ctx *ctx = nullptr;
    if (clb_ctx_create(0, &ctx) != CLB_OK) return die(ctx, "clb_ctx_create");
    int id = -1;
    if (clb_case_create(ctx, &p, geo.data(), nullptr, &id) != CLB_OK) return die(ctx, "clb_case_create");
    clb_driver::Report rep;
    if (clb_driver::setup(ctx, id, p, fs, ts, rep) != CLB_OK) return die(ctx, "set-up");
    int done = 0;
    if (steps > 0 && clb_run_steps(ctx, id, steps, &done) != CLB_OK) return die(ctx, "clb_run_steps");
    std::printf("iters");
    for (int it : rep.iters) std::printf(" %d", it);
    std::printf("\ndelta_p %.17g\ndt %.17g\nsteps %d\n", rep.delta_p, rep.dt, done);
    std::ofstream out(argv[4], std::ios::binary);
    std::vector<double> field(N);
    for (int s = 0; s < p.n_species; ++s) {
        if (clb_read_field(ctx, id, CLB_FIELD_CONC, s, field.data()) != CLB_OK) return die(ctx, "clb_read_field");
        out.write(reinterpret_cast<const char *>(field.data()), (std::streamsize) (N * sizeof(double)));
    }
    clb_case_destroy(ctx, id);
    clb_ctx_destroy(ctx);
    return 0;
}
