"""The percolation stop and the single ordered age sweep on the device,
against the oracle (FP64: fields 1e-10, counters and ages exact)."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import params as P
import oracle as O

from test_oracle_clog import sealed_channel, run_case
from test_oracle_physics import gauss_seidel_age

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("full_seal", [True, False])
def test_clog_stop_matches_oracle(ctx, full_seal):
    p, geom, fs, ts, plug = sealed_channel(full_seal)
    g = clb.Case(ctx, p, geom)
    o = O.OracleCase(p, geom)
    rg = run_case(g, p, fs, ts, plug)
    ro = run_case(o, p, fs, ts, plug)
    print("gpu", rg, "oracle", ro)
    assert rg == ro
    assert rg == ((P.CLOGGED, 1) if full_seal else (0, 12))
    assert list(g.counters()[2:8]) == list(o.counters()[2:8])
    om = o.read(P.FIELD_MASK) >= 2
    a, b = g.read(P.FIELD_CONC, 0)[om], o.read(P.FIELD_CONC, 0)[om]
    assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(b))
    assert np.array_equal(g.read(P.FIELD_MASK), o.read(P.FIELD_MASK))
    if full_seal:
        assert g.run_steps(3) == (P.CLOGGED, 0)
    g.close(); o.close()


def test_single_age_sweep_matches_oracle(ctx):
    """clb_age_update(…, 0): one in-place pass in the reference's loop order,
    solved on the device as its triangular system.  A 40x18x16 box of
    random biomass and ages, bit-exact against the oracle's in-place loop and
    against a pure-Python pass of the reference loop."""
    nx, ny, nz = 40, 18, 16
    geom = np.full((nz, ny, nx), P.PORE, np.int8)
    geom[:, 0, :] = P.WALL; geom[:, -1, :] = P.WALL; geom[0, :, :] = P.WALL; geom[-1, :, :] = P.WALL
    p = P.default_params(nx, ny, nz, 1)
    p.n_microbes = 1
    p.bmax = 10.0
    dist = np.zeros(geom.shape, int)
    inner = np.zeros(geom.shape, bool); inner[:-1, :-1, :-1] = True
    dist[(geom >= 2) & inner] = 1
    for seed in range(4):
        rng = np.random.default_rng(70 + seed)
        b = np.where(geom >= 2, rng.choice([0.0, 5.0, 10.0, 10.0, 10.0], size=geom.shape), 0.0)
        age0 = np.where(geom >= 2, rng.integers(0, 4, size=geom.shape), -1)
        g = clb.Case(ctx, p, geom)
        o = O.OracleCase(p, geom)
        for c in (g, o):
            c.write(P.FIELD_BIOMASS, 0, b)
            c.write(P.FIELD_AGE, 0, age0.astype(float))
        g.age_update(False)
        o.age_relax(False)
        ag, ao = g.read(P.FIELD_AGE).astype(int), o.read(P.FIELD_AGE).astype(int)
        assert np.array_equal(ag, ao), seed
        if seed == 0:
            assert np.array_equal(ao, gauss_seidel_age(age0, dist, b, p.bmax, once=True))
        g.age_update(True)                       # and the relaxation to the fixed point from there
        o.age_relax(True)
        assert np.array_equal(g.read(P.FIELD_AGE), o.read(P.FIELD_AGE)), seed
        g.close(); o.close()
