"""ctypes front of the CPU oracle (liboracle.so) and of the reference's own
solver/generator built into _ref/libref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None


def _params_mod():
    import sys
    pkg = os.path.join(os.path.dirname(HERE), "complab3d-gems_amd")
    if pkg not in sys.path:
        sys.path.insert(0, pkg)
    from complab_amd import params
    return params


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_case_create.restype = P
        L.orc_case_create.argtypes = [P, P, P]
        L.orc_case_destroy.argtypes = [P]
        L.orc_last_error.restype = C.c_char_p
        L.orc_last_error.argtypes = [P]
        for name, args in [
            ("orc_set_omegas", [P, P, P, C.c_double]),
            ("orc_ns_init", [P, C.c_double]),
            ("orc_ns_solve", [P, C.c_int, C.c_int, C.c_double, C.c_int, P, P]),
            ("orc_ns_couple", [P]),
            ("orc_ns_stats", [P, P]),
            ("orc_ns_update_dynamics", [P, C.c_double]),
            ("orc_ade_stabilize", [P, C.c_int]),
            ("orc_run_steps", [P, C.c_int, P]),
            ("orc_ade_steps", [P, C.c_int]),
            ("orc_equilibrium", [P]),
            ("orc_ca_spill", [P, P]),
            ("orc_ca_sweep", [P]),
            ("orc_precip_convert", [P]),
            ("orc_age_relax", [P, C.c_int, P]),
            ("orc_params_layout", [P]),
            ("orc_tracer_stop", [P, C.c_int, C.c_double, P]),
            ("orc_read_field", [P, C.c_int, C.c_int, P]),
            ("orc_write_field", [P, C.c_int, C.c_int, P]),
            ("orc_species_totals", [P, P]),
            ("orc_field_stats", [P, P]),
            ("orc_counters", [P, P]),
            ("orc_eq_solve", [P, P, P, P, P, P]),
            ("orc_eq_voxel", [P, P, P]),
            ("orc_rates", [P, P, P, C.c_int, P, P, P]),
            ("orc_thermo_factor", [P, P, P, P]),
            ("orc_set_symbolic", [P, C.c_int, P]),
            ("orc_rate_capture", [P, C.c_int]),
            ("orc_sym_evaluate", [P, P, C.c_double, P]),
            ("orc_rng_draws", [C.c_uint64, C.c_int, P]),
            ("orc_set_threads", [C.c_int]),
            ("orc_geom_spheres", [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_uint64, C.c_int, P]),
            ("orc_geom_pack", [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64, C.c_int, P, P]),
        ]:
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = args
        _lib = L
    return _lib


def ref():
    """The reference's equilibrium solver built against the Palabos stub of its
    own tests (libref.so), or None when it was not built: a cross-check."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        P = C.c_void_p
        L.ref_eq_solve.argtypes = [C.c_int, C.c_int, P, P, P, C.c_double, C.c_int, C.c_int, P, P, P, P, P]
        L.ref_eq_voxel.argtypes = [C.c_int, C.c_int, P, P, P, C.c_double, C.c_int, C.c_int, P, P]
        _ref = L
    return _ref


_ref_geom = None


def ref_geom():
    """The reference's geometry generator compiled from its own source with
    nothing standing in for anything (libref_geom.so), or None."""
    global _ref_geom
    if _ref_geom is None:
        path = os.path.join(HERE, "_ref", "libref_geom.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        P = C.c_void_p
        L.ref_rng_draws.argtypes = [C.c_uint64, C.c_int, P]
        L.ref_geom_generate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                        C.c_uint64, C.c_char_p, P]
        L.ref_geom_inspect.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, P, P, P]
        _ref_geom = L
    return _ref_geom


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleCase:
    """One case in the CPU oracle; mirrors complab_amd.Case method by method."""

    def __init__(self, params, geometry, biomass0=None):
        self.p = params
        self.shape = (params.nz, params.ny, params.nx)
        self.N = params.nx * params.ny * params.nz
        g = np.ascontiguousarray(geometry, dtype=np.int8).reshape(-1)
        assert g.size == self.N
        b = None
        if biomass0 is not None:
            b = np.ascontiguousarray(biomass0, dtype=np.float64).reshape(-1)
        self._g = g
        self.h = lib().orc_case_create(C.byref(params), _ptr(g), _ptr(b) if b is not None else None)

    def close(self):
        if self.h:
            lib().orc_case_destroy(self.h)
            self.h = None

    __del__ = close

    def _chk(self, rc):
        if rc < 0:
            raise RuntimeError("oracle error %d: %s" % (rc, lib().orc_last_error(self.h).decode()))
        return rc

    def rate_capture(self, on=True):
        self._chk(lib().orc_rate_capture(self.h, 1 if on else 0))

    def set_symbolic(self, microbe, prog):
        self._chk(lib().orc_set_symbolic(self.h, int(microbe), C.byref(prog) if prog is not None else None))

    def set_omegas(self, omp, omb, dt):
        a = np.ascontiguousarray(omp, dtype=np.float64)
        b = np.ascontiguousarray(omb, dtype=np.float64)
        self._chk(lib().orc_set_omegas(self.h, _ptr(a), _ptr(b), dt))

    def ns_init(self, delta_p):
        self._chk(lib().orc_ns_init(self.h, delta_p))

    def ns_solve(self, max_iter, rule, eps, reset=False):
        it = C.c_int(); cv = C.c_int()
        self._chk(lib().orc_ns_solve(self.h, max_iter, rule, eps, int(reset), C.byref(it), C.byref(cv)))
        return it.value, bool(cv.value)

    def ns_couple(self):
        self._chk(lib().orc_ns_couple(self.h))

    def ns_stats(self):
        out = np.zeros(4)
        self._chk(lib().orc_ns_stats(self.h, _ptr(out)))
        return out

    def ns_update_dynamics(self, perm_ratio):
        self._chk(lib().orc_ns_update_dynamics(self.h, perm_ratio))

    def ade_stabilize(self, steps):
        self._chk(lib().orc_ade_stabilize(self.h, steps))

    def run_steps(self, n):
        d = C.c_int()
        rc = lib().orc_run_steps(self.h, n, C.byref(d))
        return rc, d.value

    def ade_steps(self, n):
        self._chk(lib().orc_ade_steps(self.h, n))

    def equilibrium(self):
        self._chk(lib().orc_equilibrium(self.h))

    def ca_sweep(self):
        """One push + gather sweep (no mask/total update)."""
        self._chk(lib().orc_ca_sweep(self.h))

    def ca_spill(self):
        sw = C.c_int()
        rc = lib().orc_ca_spill(self.h, C.byref(sw))
        return rc, sw.value

    def precip_convert(self):
        """precipNodeConversion3D over the reaction box; True if any voxel converted."""
        return bool(lib().orc_precip_convert(self.h))

    def age_relax(self, to_fixed_point=True):
        k = C.c_int()
        self._chk(lib().orc_age_relax(self.h, int(to_fixed_point), C.byref(k)))
        return k.value

    def read(self, field, index=0):
        out = np.empty(self.N)
        self._chk(lib().orc_read_field(self.h, field, index, _ptr(out)))
        return out.reshape(self.shape)

    def write(self, field, index, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self._chk(lib().orc_write_field(self.h, field, index, _ptr(a)))

    def species_totals(self):
        out = np.zeros(self.p.n_species)
        self._chk(lib().orc_species_totals(self.h, _ptr(out)))
        return out

    def field_stats(self):
        """Rows (species, then microbes) of total, minimum, maximum, held."""
        out = np.zeros(4 * (self.p.n_species + self.p.n_microbes))
        self._chk(lib().orc_field_stats(self.h, _ptr(out)))
        return out.reshape(-1, 4)

    def counters(self):
        out = np.zeros(12, dtype=np.int64)
        self._chk(lib().orc_counters(self.h, _ptr(out)))
        return out


def set_threads(n=0):
    """Threads of the oracle's per-cell loops (n <= 0 only reports the count)."""
    return int(lib().orc_set_threads(int(n)))


def eq_solve(params, conc):
    c = np.ascontiguousarray(conc, dtype=np.float64)
    out = np.zeros(params.n_species)
    it = C.c_int(); res = C.c_double(); cv = C.c_int()
    lib().orc_eq_solve(C.byref(params), _ptr(c), _ptr(out), C.byref(it), C.byref(res), C.byref(cv))
    return out, it.value, res.value, bool(cv.value)


def eq_voxel(params, conc):
    c = np.ascontiguousarray(conc, dtype=np.float64)
    out = np.zeros(params.n_species)
    lib().orc_eq_voxel(C.byref(params), _ptr(c), _ptr(out))
    return out


def rates(params, conc, bio, mask):
    c = np.ascontiguousarray(conc, dtype=np.float64)
    b = np.ascontiguousarray(bio, dtype=np.float64)
    sr = np.zeros(params.n_species); br = np.zeros(max(1, params.n_microbes)); ar = np.zeros(params.n_species)
    lib().orc_rates(C.byref(params), _ptr(c), _ptr(b), mask, _ptr(sr), _ptr(br), _ptr(ar))
    return sr, br, ar


def sym_evaluate(prog, avail, bmass):
    """Outputs of one symbolic law (a params.SymProgram) at one composition."""
    a = np.zeros(64)
    a[:len(avail)] = avail
    out = np.zeros(16)
    lib().orc_sym_evaluate(C.byref(prog), _ptr(a), float(bmass), _ptr(out))
    return out[:prog.n_rates].copy()


def thermo_factor(params, conc):
    """(F_T, dG) of one composition under the parameter block's gate."""
    c = np.zeros(max(params.n_species, 1))
    c[:len(conc)] = conc
    F = C.c_double(0.0); dG = C.c_double(0.0)
    lib().orc_thermo_factor(C.byref(params), _ptr(c), C.byref(F), C.byref(dG))
    return F.value, dG.value


def rng_draws(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_rng_draws(seed, n, _ptr(out))
    return out


def geom_spheres(nx, ny, nz, radius, count, seed, walls_yz):
    out = np.zeros(nx * ny * nz, dtype=np.int8)
    lib().orc_geom_spheres(nx, ny, nz, radius, count, seed, int(walls_yz), _ptr(out))
    return out.reshape(nz, ny, nx)


def geom_pack(nx_file, ny, nz, radius, porosity, seed, kind=0):
    out = np.zeros((nx_file + 2) * ny * nz, dtype=np.int8)
    phi = C.c_double()
    rc = lib().orc_geom_pack(nx_file, ny, nz, radius, porosity, seed, kind, _ptr(out), C.byref(phi))
    return rc, out.reshape(nz, ny, nx_file + 2), phi.value


# ---- the reference's own code ----------------------------------------------------------------
def ref_eq_solve(params, conc):
    L = ref()
    ns, nc = params.n_species, params.eq_nc
    S = np.array([[params.eq_stoich[i][j] for j in range(nc)] for i in range(ns)], dtype=np.float64)
    lk = np.array([params.eq_logk[i] for i in range(ns)], dtype=np.float64)
    cs = np.array([params.eq_comp_species[j] for j in range(nc)], dtype=np.int32)
    c = np.ascontiguousarray(conc, dtype=np.float64)
    out = np.zeros(ns)
    it = C.c_int(); res = C.c_double(); cv = C.c_int()
    L.ref_eq_solve(ns, nc, _ptr(S), _ptr(lk), _ptr(cs), params.eq_tol, params.eq_max_iter,
                   params.eq_anderson_depth, _ptr(c), _ptr(out), C.byref(it), C.byref(res), C.byref(cv))
    return out, it.value, res.value, bool(cv.value)


def ref_eq_voxel(params, conc):
    L = ref()
    ns, nc = params.n_species, params.eq_nc
    S = np.array([[params.eq_stoich[i][j] for j in range(nc)] for i in range(ns)], dtype=np.float64)
    lk = np.array([params.eq_logk[i] for i in range(ns)], dtype=np.float64)
    cs = np.array([params.eq_comp_species[j] for j in range(nc)], dtype=np.int32)
    c = np.ascontiguousarray(conc, dtype=np.float64)
    out = np.zeros(ns)
    L.ref_eq_voxel(ns, nc, _ptr(S), _ptr(lk), _ptr(cs), params.eq_tol, params.eq_max_iter,
                   params.eq_anderson_depth, _ptr(c), _ptr(out))
    return out


def ref_rng_draws(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    ref_geom().ref_rng_draws(seed, n, _ptr(out))
    return out


def ref_geom_generate(nx, ny, nz, kind, radius, count, porosity, seed, walls):
    out = np.zeros(nx * ny * nz, dtype=np.int8)
    rc = ref_geom().ref_geom_generate(nx, ny, nz, kind, radius, count, porosity, seed, walls.encode(), _ptr(out))
    assert rc == 0
    return out.reshape(nz, ny, nx)


def ref_geom_inspect(codes):
    nz, ny, nx = codes.shape
    a = np.ascontiguousarray(codes, dtype=np.int8)
    phi = C.c_double(); op = C.c_long(); cl = C.c_int(); sp = C.c_int(); pc = C.c_int()
    ref_geom().ref_geom_inspect(nx, ny, nz, _ptr(a), C.byref(phi), C.byref(op), C.byref(cl), C.byref(sp), C.byref(pc))
    return dict(porosity=phi.value, open=op.value, clusters=cl.value, spanning=sp.value, percolates=bool(pc.value))
