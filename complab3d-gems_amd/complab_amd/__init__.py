"""Python front of libclb.so, the MI355X-native coupled reaction step.

Everything here is plumbing over the C-ABI of include/clb.h: the device work
is done by the HIP kernels in ../csrc.  Importing this module does not touch
the GPU; `Context()` does, and raises if the HIP library is missing or no
device is present -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import params as P
from .params import CaseParams, default_params, copy_params  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libclb.so")
_lib = None


class ClbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("clb error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Loads libclb.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libclb.so is not built (run __graft_entry__.build()); "
                           "the HIP extension is required, there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    V = C.c_void_p
    I = C.c_int
    D = C.c_double
    sig = {
        "clb_abi_version": ([], I),
        "clb_ctx_create": ([I, C.POINTER(V)], I),
        "clb_ctx_destroy": ([V], I),
        "clb_last_error": ([V], C.c_char_p),
        "clb_device_name": ([V, C.c_char_p, C.c_size_t], I),
        "clb_sync": ([V], I),
        "clb_case_create": ([V, V, V, V, C.POINTER(I)], I),
        "clb_case_destroy": ([V, I], I),
        "clb_case_footprint": ([V], C.c_size_t),
        "clb_case_get_params": ([V, I, V], I),
        "clb_ns_init": ([V, I, D], I),
        "clb_ns_solve": ([V, I, I, I, D, I, C.POINTER(I), C.POINTER(I)], I),
        "clb_ns_couple": ([V, I], I),
        "clb_ns_stats": ([V, I, V], I),
        "clb_ns_update_dynamics": ([V, I, D], I),
        "clb_set_ade_relaxation": ([V, I, V, V, D], I),
        "clb_ade_stabilize": ([V, I, I], I),
        "clb_run_steps": ([V, I, I, C.POINTER(I)], I),
        "clb_ade_steps": ([V, I, I], I),
        "clb_equilibrium": ([V, I], I),
        "clb_ca_spill": ([V, I, C.POINTER(I)], I),
        "clb_ade_collide_stream": ([V, I, I], I),
        "clb_kinetics": ([V, I, I], I),
        "clb_mask_total": ([V, I, C.POINTER(D), C.POINTER(C.c_int64)], I),
        "clb_ca_sweep": ([V, I], I),
        "clb_age_update": ([V, I, I], I),
        "clb_update_solute_dynamics": ([V, I], I),
        "clb_update_solute_solid": ([V, I, C.POINTER(I)], I),
        "clb_precip_convert": ([V, I, C.POINTER(I)], I),
        "clb_dissol_convert": ([V, I, C.POINTER(I)], I),
        "clb_read_field": ([V, I, I, I, V], I),
        "clb_write_field": ([V, I, I, I, V], I),
        "clb_species_totals": ([V, I, V], I),
        "clb_species_totals_noalloc": ([V, I, V], I),
        "clb_set_tuning": ([V, I, I, I], I),
        "clb_field_stats": ([V, I, V], I),
        "clb_counters": ([V, I, V], I),
        "clb_kernel_time": ([V, I, I, C.POINTER(D), C.POINTER(C.c_int64)], I),
        "clb_set_timing_stride": ([V, I, I], I),
        "clb_set_symbolic": ([V, I, I, V], I),
        "clb_rate_capture": ([V, I, I], I),
        "clb_checkpoint_size": ([V, I, C.POINTER(C.c_size_t)], I),
        "clb_checkpoint_save": ([V, I, V, C.c_size_t], I),
        "clb_checkpoint_load": ([V, I, V, C.c_size_t], I),
        "clb_stream_copy": ([V, C.c_size_t, I, C.POINTER(D), C.POINTER(D)], I),
        "clb_rng_draws": ([C.c_uint64, I, V], I),
        "clb_geom_spheres": ([I, I, I, D, I, C.c_uint64, I, V], I),
        "clb_geom_pack": ([I, I, I, D, D, C.c_uint64, I, V, C.POINTER(D)], I),
        "clb_geom_generate": ([I, I, I, I, D, I, D, D, D, C.c_uint64, I, V], I),
        "clb_batch_create": ([V, I, V, I, C.POINTER(I)], I),
        "clb_batch_destroy": ([V, I], I),
        "clb_batch_add": ([V, I, I], I),
        "clb_batch_remove": ([V, I, I], I),
        "clb_batch_members": ([V, I, C.POINTER(I), V, I], I),
        "clb_batch_ns_solve": ([V, I, I, I, D, I, V, V, V], I),
        "clb_batch_ade_stabilize": ([V, I, I, V], I),
        "clb_batch_run_steps": ([V, I, I, V, V], I),
        "clb_batch_set_tile": ([V, I, I, I], I),
        "clb_batch_tuning": ([V, I, I, C.POINTER(I), V], I),
        "clb_flow_kernel_time": ([V, I, I, C.POINTER(D), C.POINTER(C.c_int64)], I),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    if L.clb_abi_version() != P.ABI_VERSION:
        raise RuntimeError("libclb.so has ABI version %d, this package expects %d: rebuild it"
                           % (L.clb_abi_version(), P.ABI_VERSION))
    _lib = L
    return L


def exported_symbols():
    """Names of the entry points declared in include/clb.h."""
    return [
        "clb_abi_version", "clb_ctx_create", "clb_ctx_destroy", "clb_last_error", "clb_device_name", "clb_sync",
        "clb_case_create", "clb_case_destroy", "clb_case_footprint", "clb_case_get_params", "clb_ns_init", "clb_ns_solve",
        "clb_ns_couple", "clb_ns_stats", "clb_ns_update_dynamics", "clb_set_ade_relaxation",
        "clb_ade_stabilize", "clb_run_steps", "clb_ade_steps", "clb_equilibrium", "clb_ca_spill",
        "clb_ade_collide_stream", "clb_kinetics", "clb_mask_total", "clb_ca_sweep", "clb_age_update",
        "clb_update_solute_dynamics", "clb_update_solute_solid", "clb_precip_convert", "clb_dissol_convert",
        "clb_read_field", "clb_write_field", "clb_species_totals", "clb_field_stats", "clb_counters", "clb_kernel_time", "clb_set_timing_stride",
        "clb_set_symbolic", "clb_rate_capture", "clb_checkpoint_size", "clb_checkpoint_save", "clb_checkpoint_load",
        "clb_stream_copy",
        "clb_rng_draws", "clb_geom_spheres", "clb_geom_pack", "clb_geom_generate",
        "clb_species_totals_noalloc", "clb_set_tuning",
        "clb_batch_create", "clb_batch_destroy", "clb_batch_add", "clb_batch_remove", "clb_batch_members",
        "clb_batch_ns_solve", "clb_batch_ade_stabilize", "clb_batch_run_steps", "clb_batch_set_tile",
        "clb_batch_tuning", "clb_flow_kernel_time",
    ]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One device, one HIP stream (clb_ctx)."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.clb_ctx_create(device, C.byref(h))
        if rc != 0:
            raise ClbError(rc, "no usable HIP device %d" % device)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().clb_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc < 0:
            raise ClbError(rc, lib().clb_last_error(self.h).decode())
        return rc

    def device_name(self):
        buf = C.create_string_buffer(256)
        self.check(lib().clb_device_name(self.h, buf, 256))
        return buf.value.decode()

    def sync(self):
        self.check(lib().clb_sync(self.h))

    def stream_copy(self, nbytes, reps=20):
        """Plain device copy bandwidth (GB/s, read + write bytes) and best time (ms)."""
        g = C.c_double(); ms = C.c_double()
        self.check(lib().clb_stream_copy(self.h, nbytes, reps, C.byref(g), C.byref(ms)))
        return g.value, ms.value


class Case:
    """One simulation on the device; the same methods as the oracle's case."""

    def __init__(self, ctx, params, geometry, biomass0=None):
        self.ctx = ctx
        self.p = params
        self.shape = (params.nz, params.ny, params.nx)
        self.N = params.nx * params.ny * params.nz
        g = np.ascontiguousarray(geometry, dtype=np.int8).reshape(-1)
        if g.size != self.N:
            raise ValueError("geometry has %d voxels, the lattice %d" % (g.size, self.N))
        b = None if biomass0 is None else np.ascontiguousarray(biomass0, dtype=np.float64).reshape(-1)
        cid = C.c_int()
        ctx.check(lib().clb_case_create(ctx.h, C.byref(params), _p(g), _p(b) if b is not None else None,
                                        C.byref(cid)))
        self.id = cid.value

    def close(self):
        if getattr(self, "id", None) is not None and self.ctx.h:
            lib().clb_case_destroy(self.ctx.h, self.id)
            self.id = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        return self.ctx.check(rc)

    def get_params(self):
        out = P.CaseParams()
        self._c(lib().clb_case_get_params(self.ctx.h, self.id, C.byref(out)))
        return out

    def set_omegas(self, omp, omb, dt):
        a = np.ascontiguousarray(omp, dtype=np.float64)
        b = np.ascontiguousarray(omb, dtype=np.float64)
        self._c(lib().clb_set_ade_relaxation(self.ctx.h, self.id, _p(a), _p(b), dt))

    def ns_init(self, delta_p):
        self._c(lib().clb_ns_init(self.ctx.h, self.id, delta_p))

    def ns_solve(self, max_iter, rule, eps, reset=False):
        it = C.c_int(); cv = C.c_int()
        self._c(lib().clb_ns_solve(self.ctx.h, self.id, max_iter, rule, eps, int(reset), C.byref(it), C.byref(cv)))
        return it.value, bool(cv.value)

    def ns_couple(self):
        self._c(lib().clb_ns_couple(self.ctx.h, self.id))

    def ns_stats(self):
        out = np.zeros(4)
        self._c(lib().clb_ns_stats(self.ctx.h, self.id, _p(out)))
        return out

    def ns_update_dynamics(self, perm_ratio):
        self._c(lib().clb_ns_update_dynamics(self.ctx.h, self.id, perm_ratio))

    def ade_stabilize(self, steps):
        self._c(lib().clb_ade_stabilize(self.ctx.h, self.id, steps))

    def run_steps(self, n):
        d = C.c_int()
        rc = lib().clb_run_steps(self.ctx.h, self.id, n, C.byref(d))
        return rc, d.value

    def ade_steps(self, n):
        self._c(lib().clb_ade_steps(self.ctx.h, self.id, n))

    def equilibrium(self):
        self._c(lib().clb_equilibrium(self.ctx.h, self.id))

    def ca_spill(self):
        sw = C.c_int()
        rc = lib().clb_ca_spill(self.ctx.h, self.id, C.byref(sw))
        return rc, sw.value

    # ---- the processors one at a time
    def ade_collide_stream(self, n):
        self._c(lib().clb_ade_collide_stream(self.ctx.h, self.id, n))

    def kinetics(self, which=P.KIN_BIOTIC | P.KIN_ABIOTIC):
        self._c(lib().clb_kinetics(self.ctx.h, self.id, which))

    def mask_total(self):
        g = C.c_double(); m = C.c_int64()
        self._c(lib().clb_mask_total(self.ctx.h, self.id, C.byref(g), C.byref(m)))
        return g.value, m.value

    def ca_sweep(self):
        self._c(lib().clb_ca_sweep(self.ctx.h, self.id))

    def age_update(self, to_fixed_point=True):
        self._c(lib().clb_age_update(self.ctx.h, self.id, int(to_fixed_point)))

    def update_solute_dynamics(self):
        self._c(lib().clb_update_solute_dynamics(self.ctx.h, self.id))

    def update_solute_solid(self):
        n = C.c_int()
        self._c(lib().clb_update_solute_solid(self.ctx.h, self.id, C.byref(n)))
        return n.value

    def precip_convert(self):
        ch = C.c_int()
        self._c(lib().clb_precip_convert(self.ctx.h, self.id, C.byref(ch)))
        return bool(ch.value)

    def dissol_convert(self):
        n = C.c_int()
        self._c(lib().clb_dissol_convert(self.ctx.h, self.id, C.byref(n)))
        return n.value

    def read(self, field, index=0):
        out = np.empty(self.N)
        self._c(lib().clb_read_field(self.ctx.h, self.id, field, index, _p(out)))
        return out.reshape(self.shape)

    def write(self, field, index, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self._c(lib().clb_write_field(self.ctx.h, self.id, field, index, _p(a)))

    def species_totals(self):
        out = np.zeros(self.p.n_species)
        self._c(lib().clb_species_totals(self.ctx.h, self.id, _p(out)))
        return out

    def field_stats(self):
        """Rows (species, then microbes) of total, minimum, maximum, held."""
        out = np.zeros(4 * (self.p.n_species + self.p.n_microbes))
        self._c(lib().clb_field_stats(self.ctx.h, self.id, _p(out)))
        return out.reshape(-1, 4)

    def counters(self):
        out = np.zeros(12, dtype=np.int64)
        self._c(lib().clb_counters(self.ctx.h, self.id, _p(out)))
        return out

    def set_symbolic(self, microbe, prog):
        """Microbe `microbe` takes its biotic rates from `prog` (a
        params.SymProgram, from complab_amd.symbolic.compile_law); None
        returns it to the compiled law.  microbe = params.SYM_ABIOTIC sets
        (or, with None, clears) the abiotic law."""
        self._c(lib().clb_set_symbolic(self.ctx.h, self.id, int(microbe),
                                       C.byref(prog) if prog is not None else None))

    def rate_capture(self, on=True):
        """Record each following step's requested reaction increments / dt
        (read as params.FIELD_RATE; clb_rate_capture)."""
        self._c(lib().clb_rate_capture(self.ctx.h, self.id, 1 if on else 0))

    def checkpoint(self) -> bytes:
        """The case's state as one block (clb_checkpoint_save)."""
        n = C.c_size_t(0)
        self._c(lib().clb_checkpoint_size(self.ctx.h, self.id, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._c(lib().clb_checkpoint_save(self.ctx.h, self.id, buf, n))
        return buf.raw

    def restore(self, data: bytes):
        """Loads a block from checkpoint() of a case with the same parameters
        and geometry; the case then continues bit-identically."""
        buf = C.create_string_buffer(bytes(data), len(data))
        self._c(lib().clb_checkpoint_load(self.ctx.h, self.id, buf, C.c_size_t(len(data))))

    def save_checkpoint(self, path):
        tmp = path + ".part"
        with open(tmp, "wb") as f:
            f.write(self.checkpoint())
        os.replace(tmp, path)

    def load_checkpoint(self, path):
        with open(path, "rb") as f:
            self.restore(f.read())

    def set_tuning(self, knob, value):
        """A knob of the hot kernels (P.TUNE_*: the flow launch shape, the
        transport bounce-back store before the first transport launch);
        results do not change."""
        self._c(lib().clb_set_tuning(self.ctx.h, self.id, int(knob), int(value)))

    def set_timing_stride(self, stride):
        """Time only every stride-th transport launch (event markers cost the
        device a dispatch gap each)."""
        self._c(lib().clb_set_timing_stride(self.ctx.h, self.id, int(stride)))

    def kernel_time(self, reset=False):
        ms = C.c_double(); n = C.c_int64()
        self._c(lib().clb_kernel_time(self.ctx.h, self.id, int(reset), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def flow_kernel_time(self, reset=False):
        """Average device time (ms) of the timed flow iterations (k_ns_step
        alone) and their count."""
        ms = C.c_double(); n = C.c_int64()
        self._c(lib().clb_flow_kernel_time(self.ctx.h, self.id, int(reset), C.byref(ms), C.byref(n)))
        return ms.value, n.value


def footprint(params):
    return lib().clb_case_footprint(C.byref(params))


def geom_pack(nx_file, ny, nz, radius, porosity, seed, kind=P.GEOM_SPHERES):
    out = np.zeros((nx_file + 2) * ny * nz, dtype=np.int8)
    phi = C.c_double()
    rc = lib().clb_geom_pack(nx_file, ny, nz, radius, porosity, seed, kind, _p(out), C.byref(phi))
    return rc, out.reshape(nz, ny, nx_file + 2), phi.value


def geom_spheres(nx, ny, nz, radius, count, seed, walls_yz):
    out = np.zeros(nx * ny * nz, dtype=np.int8)
    rc = lib().clb_geom_spheres(nx, ny, nz, radius, count, seed, int(walls_yz), _p(out))
    if rc:
        raise ClbError(rc, "clb_geom_spheres")
    return out.reshape(nz, ny, nx)


def geom_generate(nx, ny, nz, kind, radius=6.0, count=20, porosity=0.4, aperture=6.0, roughness=2.0,
                  seed=1, walls=2):
    """complab_geom::generate of the reference: (nz, ny, nx) material codes."""
    out = np.zeros(nx * ny * nz, dtype=np.int8)
    rc = lib().clb_geom_generate(nx, ny, nz, kind, radius, count, porosity, aperture, roughness, seed, walls, _p(out))
    if rc:
        raise ClbError(rc, "clb_geom_generate")
    return out.reshape(nz, ny, nx)


def rng_draws(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    lib().clb_rng_draws(seed, n, _p(out))
    return out


class Batch:
    """Cases of one precision stepped together (clb_batch_*): each kernel of a
    flow iteration or a transport step covers every member, and every member
    ends where its own calls would take it.  Per-member results come back in
    member order (self.cases)."""

    def __init__(self, ctx, cases=(), precision=None):
        cases = list(cases)
        if precision is None:
            if not cases:
                raise ValueError("an empty batch needs its precision")
            precision = cases[0].p.precision
        self.ctx = ctx
        self.precision = precision
        ids = np.array([c.id for c in cases], dtype=np.int32)
        bid = C.c_int()
        ctx.check(lib().clb_batch_create(ctx.h, precision, _p(ids) if len(ids) else None, len(ids), C.byref(bid)))
        self.id = bid.value
        self.cases = cases

    def close(self):
        if getattr(self, "id", None) is not None and self.ctx.h:
            lib().clb_batch_destroy(self.ctx.h, self.id)
            self.id = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.cases)

    def add(self, case):
        self.ctx.check(lib().clb_batch_add(self.ctx.h, self.id, case.id))
        self.cases.append(case)

    def remove(self, case):
        self.ctx.check(lib().clb_batch_remove(self.ctx.h, self.id, case.id))
        self.cases = [c for c in self.cases if c is not case]

    def ns_solve(self, max_iter, rule, eps, reset=False):
        """Per member (iterations, converged, status) of clb_ns_solve."""
        n = len(self.cases)
        it = np.zeros(n, np.int32); cv = np.zeros(n, np.int32); st = np.zeros(n, np.int32)
        self.ctx.check(lib().clb_batch_ns_solve(self.ctx.h, self.id, max_iter, rule, eps, int(reset),
                                                _p(it), _p(cv), _p(st)))
        return [(int(a), bool(b), int(c)) for a, b, c in zip(it, cv, st)]

    def ade_stabilize(self, steps):
        st = np.zeros(len(self.cases), np.int32)
        self.ctx.check(lib().clb_batch_ade_stabilize(self.ctx.h, self.id, steps, _p(st)))
        return [int(x) for x in st]

    def run_steps(self, n):
        """Per member (status, steps done) of clb_run_steps."""
        k = len(self.cases)
        d = np.zeros(k, np.int32); st = np.zeros(k, np.int32)
        self.ctx.check(lib().clb_batch_run_steps(self.ctx.h, self.id, n, _p(d), _p(st)))
        return [(int(a), int(b)) for a, b in zip(st, d)]

    def set_tile(self, mode, tile):
        self.ctx.check(lib().clb_batch_set_tile(self.ctx.h, self.id, mode, tile))

    def tuning(self, mode=P.BATCH_STEPS):
        """(picked workgroup size or 0, {candidate: microseconds per step})."""
        t = C.c_int(); us = np.zeros(3)
        self.ctx.check(lib().clb_batch_tuning(self.ctx.h, self.id, mode, C.byref(t), _p(us)))
        return t.value, {256: float(us[0]), 128: float(us[1]), 64: float(us[2])}
