"""clb_species_totals_noalloc (the entry point that allocates nothing) against
clb_species_totals through the C-ABI: two calls in a row on two cases, the
second smaller, equal bit for bit (the same kernels in the same order, only
the scratch differs); a case whose reductions exceed the static scratch is
refused with CLB_E_ARG."""
import ctypes as C

import numpy as np
import pytest

import complab_amd as clb
from complab_amd import configs
from complab_amd import params as P

pytestmark = pytest.mark.gpu


def totals_both(case):
    L = clb.lib()
    a = np.zeros(case.p.n_species)
    b = np.full(case.p.n_species, np.nan)
    assert L.clb_species_totals(case.ctx.h, case.id, a.ctypes.data_as(C.c_void_p)) == 0
    assert L.clb_species_totals_noalloc(case.ctx.h, case.id, b.ctypes.data_as(C.c_void_p)) == 0
    return a, b


def mixed_case(ctx, scale):
    w = configs.config1(scale=scale)
    p = w.params
    p.n_species = 3
    p.mobile[2] = 0                                    # an immobile species takes the dense reduction
    p.c0[2] = 0.7
    p.left_btype[2] = P.BC_CLOSED; p.right_btype[2] = P.BC_CLOSED
    g = clb.Case(ctx, p, w.geometry)
    rng = np.random.default_rng(int(scale * 100))
    g.write(P.FIELD_CONC, 0, rng.uniform(0, 1, w.geometry.shape))
    g.write(P.FIELD_CONC, 2, rng.uniform(0, 2, w.geometry.shape))
    g.set_omegas([1.2, 1.1, 1.0], [1.2, 1.1, 1.0], 1.0)
    g.ade_steps(5)
    return g


def test_noalloc_totals_equal_the_allocating_entry_point(ctx):
    big = mixed_case(ctx, 1.0)
    small = mixed_case(ctx, 0.25)
    assert big.p.nx * big.p.ny * big.p.nz > small.p.nx * small.p.ny * small.p.nz
    for case in (big, small):                          # the second call on a smaller input
        a, b = totals_both(case)
        print(case.p.nx, case.p.ny, case.p.nz, a, b)
        assert np.all(np.isfinite(a)) and np.all(a > 0)
        assert np.array_equal(a, b)
    a, b = totals_both(big)                            # and back: nothing stale from the small call
    assert np.array_equal(a, b)
    big.close(); small.close()


def test_noalloc_refuses_a_case_beyond_its_scratch(ctx):
    """An immobile species needs 4 partials per 256 voxels: 42.5 M voxels
    exceed the 5 MiB array.  The lattice is all bounce-back but for one
    channel, so the case is cheap to build."""
    nx, ny, nz = 1000, 206, 206
    geom = np.full((nz, ny, nx), P.WALL, np.int8)
    geom[100:103, 100:103, :] = P.PORE
    p = P.default_params(nx, ny, nz, 2)
    p.mobile[1] = 0
    g = clb.Case(ctx, p, geom)
    out = np.zeros(2)
    L = clb.lib()
    rc = L.clb_species_totals_noalloc(ctx.h, g.id, out.ctypes.data_as(C.c_void_p))
    assert rc == P.E_ARG
    assert b"scratch" in L.clb_last_error(ctx.h)
    a = g.species_totals()                             # the allocating entry point still serves it
    assert np.all(np.isfinite(a))
    g.close()
