/*
 * ref_driver.cpp -- TEST INFRASTRUCTURE.  Builds the reference's equilibrium
 * solver (src/complab3d_processors_part4_eqsolver.hh) against the Palabos
 * stub the reference's own tests use (tests/palabos_stub.hh), as its
 * tests/run_tests.sh compiles its processors, and exposes it through
 * extern "C".  A build against a stand-in for Palabos is a cross-check of the
 * oracle's solver, not a pin of it: the oracle's speciation is pinned by the
 * mathematics (closed forms, mass balance and mass action at the stated
 * tolerances, tests/test_oracle_golden.py).  Built into oracle/_ref/ by
 * oracle/Makefile from the sources where they lie; nothing of the reference is
 * copied into this repository.
 */
#include "palabos_stub.hh"

namespace plb { PlbOut pcout; }

#include "complab3d_processors_part4_eqsolver.hh"

#include <cstdint>
#include <string>
#include <vector>

namespace {

/* species are named S0..S(ns-1); component j takes the name of the species
 * that carries its initial guess (or a name no species has, so the solver's
 * default guess 1e-7 applies, part4_eqsolver.hh:987-1002) */
EquilibriumChemistry<double> makeSolver(int ns, int nc, const double *stoich /* ns*nc */,
                                        const double *logk, const int *compSpecies,
                                        double tol, int maxIter, int depth) {
    EquilibriumChemistry<double> eq;
    std::vector<std::string> sn, cn;
    for (int i = 0; i < ns; ++i) sn.push_back("S" + std::to_string(i));
    for (int j = 0; j < nc; ++j) cn.push_back(compSpecies[j] >= 0 ? "S" + std::to_string(compSpecies[j]) : "COMP" + std::to_string(j));
    std::vector<std::vector<double>> S(ns, std::vector<double>(nc));
    for (int i = 0; i < ns; ++i) for (int j = 0; j < nc; ++j) S[i][j] = stoich[i * nc + j];
    eq.setSpeciesNames(sn);
    eq.setComponentNames(cn);
    eq.setLogK(std::vector<double>(logk, logk + ns));
    eq.setStoichiometryMatrix(S);
    eq.setTolerance(tol);
    eq.setMaxIterations(maxIter);
    eq.setAndersonDepth(depth);
    return eq;
}

}  // namespace

extern "C" {

/* EquilibriumChemistry<double>::calculate_species_concentrations */
int ref_eq_solve(int ns, int nc, const double *stoich, const double *logk, const int *compSpecies,
                 double tol, int maxIter, int depth, const double *cin, double *cout,
                 int *iters, double *residual, int *converged) {
    EquilibriumChemistry<double> eq = makeSolver(ns, nc, stoich, logk, compSpecies, tol, maxIter, depth);
    std::vector<double> c(cin, cin + ns);
    std::vector<double> r = eq.calculate_species_concentrations(c);
    for (int i = 0; i < ns; ++i) cout[i] = r[i];
    if (iters) *iters = (int) eq.getLastIterations();
    if (residual) *residual = eq.getLastResidual();
    if (converged) *converged = eq.didConverge() ? 1 : 0;
    return 0;
}

/* run_equilibrium_biotic::process on one voxel of the stub lattice: density
 * in = conc, out = conc + the limited increment it deposits */
int ref_eq_voxel(int ns, int nc, const double *stoich, const double *logk, const int *compSpecies,
                 double tol, int maxIter, int depth, const double *cin, double *cout) {
    EquilibriumChemistry<double> eq = makeSolver(ns, nc, stoich, logk, compSpecies, tol, maxIter, depth);
    std::vector<BlockLattice3D<double, D3Q7>> lat((size_t) ns + 1);
    std::vector<BlockLattice3D<double, D3Q7> *> ptr;
    for (int i = 0; i < ns; ++i) { lat[i].c.rho = cin[i]; ptr.push_back(&lat[i]); }
    lat[ns].c.rho = 2.0;                         /* pore */
    ptr.push_back(&lat[ns]);
    run_equilibrium_biotic<double, D3Q7> proc(3, ns, eq, 0, 1);
    Box3D box = {1, 1, 0, 0, 0, 0};
    proc.process(box, ptr);
    for (int i = 0; i < ns; ++i) cout[i] = lat[i].c.computeDensity();
    return 0;
}

}  // extern "C"
