"""Writes the golden fixtures of tests/golden from the reference's own code.

Runs in a container that holds the CompLaB3D-GEMS tree at /root/reference.
`make -C oracle ref` compiles from the sources where they lie:
  * the geometry generator (src/complab3d_geometry.hh, standard library only)
    into oracle/_ref/libref_geom.so -- the reference compiled here: the
    geometry fixtures are its outputs;
  * the equilibrium solver (src/complab3d_processors_part4_eqsolver.hh)
    against the Palabos stub of the reference's own tests into
    oracle/_ref/libref.so -- a build on a stand-in for Palabos, so the
    equilibrium fixtures are a cross-check, not a pin (the oracle's
    speciation is pinned by tests/test_oracle_eq_laws.py).
This script calls those libraries on seeded inputs and stores inputs and
outputs as .npz data.  The ABC sweep geometry fixtures are read from the
reference tree as data.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "complab3d-gems_amd"))

import oracle as O                      # noqa: E402
from complab_amd import params as P     # noqa: E402

REF_TREE = "/root/reference"


def carbonate(p):
    """examples/04_equilibrium/CompLaB.xml:145-163"""
    p.n_species = 4
    p.eq_on = 1
    p.eq_nc = 2
    S = [[1, 0], [0, 1], [1, -1], [1, 1]]
    lk = [0.0, 0.0, -10.33, 6.35]
    for i in range(4):
        for j in range(2):
            p.eq_stoich[i][j] = S[i][j]
        p.eq_logk[i] = lk[i]
    p.eq_comp_species[0] = 0
    p.eq_comp_species[1] = 1


def calcium_carbonate(p, rng=None):
    """A 6-component carbonate-calcium tableau: components H, CO3, Ca, Mg, Na,
    Cl; complexes HCO3 (H+CO3), CaCO3(aq) (Ca+CO3); log K optionally drawn
    within +-2 of nominal."""
    p.n_species = 8
    p.eq_on = 1
    p.eq_nc = 6
    rows = [[1, 0, 0, 0, 0, 0], [0, 1, 0, 0, 0, 0], [0, 0, 1, 0, 0, 0], [0, 0, 0, 1, 0, 0],
            [0, 0, 0, 0, 1, 0], [0, 0, 0, 0, 0, 1], [1, 1, 0, 0, 0, 0], [0, 1, 1, 0, 0, 0]]
    lk = [0, 0, 0, 0, 0, 0, 10.33, 3.22]
    if rng is not None:
        lk = [v + (rng.uniform(-2, 2) if i >= 6 else 0.0) for i, v in enumerate(lk)]
    for i in range(8):
        for j in range(6):
            p.eq_stoich[i][j] = rows[i][j]
        p.eq_logk[i] = lk[i]
    for j in range(6):
        p.eq_comp_species[j] = j


def eq_vectors(p, n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    cin = 10 ** rng.uniform(lo, hi, size=(n, p.n_species))
    out = np.zeros_like(cin); its = np.zeros(n, np.int32); res = np.zeros(n); cv = np.zeros(n, np.int8)
    vox = np.zeros_like(cin)
    for k in range(n):
        out[k], its[k], res[k], cv[k] = O.ref_eq_solve(p, cin[k])
        vox[k] = O.ref_eq_voxel(p, cin[k])
    return dict(cin=cin, out=out, iters=its, residual=res, converged=cv, voxel=vox)


def read_dat(path, dims):
    nx, ny, nz = dims
    v = np.loadtxt(path, dtype=np.int64)
    return v


def geometry_fracture():
    """complab_geom::generate "fracture" (aperture 6, roughness 2: GenOptions'
    defaults, complab3d_geometry.hh:113) with walls on y"""
    frac = O.ref_geom_generate(40, 24, 8, 3, 3.0, 0, 0.3, 9, "y")
    np.savez_compressed(os.path.join(HERE, "geometry_fracture.npz"), fracture=frac)


def main():
    if O.ref() is None or O.ref_geom() is None:
        raise SystemExit("oracle/_ref is missing: run `make -C oracle ref` where the reference tree exists")
    # ---- equilibrium: carbonate (2 components) and calcium carbonate (6 components)
    p = P.default_params(3, 1, 1, 4)
    carbonate(p)
    p.eq_tol = 1e-8; p.eq_max_iter = 200; p.eq_anderson_depth = 4
    np.savez_compressed(os.path.join(HERE, "eq_carbonate.npz"), **eq_vectors(p, 64, 7, -7, -1))
    p6 = P.default_params(3, 1, 1, 8)
    calcium_carbonate(p6, np.random.default_rng(3))
    p6.eq_tol = 1e-8; p6.eq_max_iter = 200; p6.eq_anderson_depth = 4
    d = eq_vectors(p6, 64, 9, -6, -2)
    d["logk"] = np.array([p6.eq_logk[i] for i in range(8)])
    np.savez_compressed(os.path.join(HERE, "eq_calcium_carbonate.npz"), **d)
    # ---- geometry: the reference generator and its Rng
    draws = O.ref_rng_draws(20240601, 32)
    sph = O.ref_geom_generate(24, 16, 12, 0, 3.0, 14, 0.4, 77, "yz")
    cyl = O.ref_geom_generate(20, 14, 6, 2, 2.5, 9, 0.4, 5, "y")
    rnd = O.ref_geom_generate(16, 12, 10, 1, 3.0, 0, 0.35, 11, "none")
    np.savez_compressed(os.path.join(HERE, "geometry_generate.npz"), rng_seed=np.uint64(20240601), rng=draws,
                        spheres=sph, cylinders=cyl, random=rnd)
    geometry_fracture()
    # ---- ABC sweep geometry fixtures: porosity / percolation of the shipped files
    gdir = os.path.join(REF_TREE, "ABC_sweep_package", "geometries")
    insp = {}
    for name in ("geom_0000", "geom_0004", "geom_0900"):
        path = os.path.join(gdir, name, "geometry.dat")
        if not os.path.exists(path):
            continue
        dims = [int(t) for t in open(os.path.join(gdir, name, "geometry.dims")).read().split()]
        v = read_dat(path, dims).astype(np.int8)
        nx, ny, nz = dims
        # the file is written with z fastest (readGeometry's Palabos slice order)
        codes = v.reshape(nx, ny, nz).transpose(2, 1, 0).copy()
        r = O.ref_geom_inspect(codes)
        insp[name] = dict(dims=dims, counts={str(k): int((codes == k).sum()) for k in (0, 1, 2)}, **r)
        if name == "geom_0000":
            np.savez_compressed(os.path.join(HERE, "abc_geom_0000.npz"), codes=codes)
    with open(os.path.join(HERE, "abc_geometries.json"), "w") as f:
        json.dump(insp, f, indent=1, sort_keys=True)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    if sys.argv[1:] == ["fracture"]:
        geometry_fracture()
        sys.exit(0)
    main()
