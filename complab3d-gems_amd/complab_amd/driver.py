"""Set-up sequence of a CompLaB3D run on any case object (GPU or oracle).

The order and formulas are those of complab.cpp: the flow solve, the
permeability and the pressure-drop correction for the target Peclet number
(complab.cpp:1111-1183), the lattice relaxation of every advection-diffusion
lattice (complab.cpp:1220-1283), the initial mask update (1981-2000), the
velocity coupling and the 10 000-step stabilization with its reset
(2009-2042).  `case` is anything with the methods of complab_amd.Case.
"""
from dataclasses import dataclass, field
from typing import List

from . import params as P

CS2_ADE = 0.25        # D3Q7: rest 1/4, faces 1/8  =>  cs^2 = 1/4
CS2_NS = 1.0 / 3.0


@dataclass
class FlowSetup:
    pe: float                       # target Peclet number (0: no flow)
    char_length: float              # characteristic length, lattice units
    delta_p: float                  # first pressure drop
    ns_tau: float = 0.8
    max_iter: int = 100000          # ns_max_iT1
    rule: int = P.CONV_TRACER
    eps: float = 1e-8               # ns_converge_iT1
    tau_ade_fixed: float = 0.8      # complab.cpp:332
    has_biofilm: bool = False       # bfilm_count > 0 (pore + biofilm mean velocity)


@dataclass
class TransportSetup:
    d_pore: List[float]             # physical pore diffusivity per species
    d_biofilm: List[float]
    dx: float = 1.0
    tau_if_no_flow: float = 0.8     # refTau = tau when Pe = 0
    stabilize_steps: int = 10000


@dataclass
class SetupReport:
    delta_p: float = 0.0
    iters: List[int] = field(default_factory=list)
    converged: List[bool] = field(default_factory=list)
    pore_mean_u: float = 0.0
    max_ux: float = 0.0
    outlet_ux: float = 0.0
    permeability: float = 0.0
    pe_achieved: float = 0.0
    ref_tau: float = 0.0
    dt: float = 0.0
    omega_pore: List[float] = field(default_factory=list)
    omega_biofilm: List[float] = field(default_factory=list)
    tau_notes: List[str] = field(default_factory=list)   # the tau screen's warnings


TAU_REJECT = 0.51     # complab.cpp:1318: below this the lattice rings
TAU_WARN = 0.55       # complab.cpp:1319


class TauError(ValueError):
    """A relaxation time below TAU_REJECT: the reference terminates (complab.cpp:1358-1377)."""


def tau_screen(p, omega_pore, omega_biofilm):
    """The relaxation times every streaming species will use, checked before
    the run as complab.cpp:1315-1378 does: the pore and biofilm tau of each
    mobile species (an immobile one never streams).  Returns the warnings
    (tau in [0.51, 0.55), or above 2); raises TauError below 0.51."""
    low, marginal, high = [], [], []
    for s in range(p.n_species):
        if not p.mobile[s]:
            continue
        for where, om in (("pore", omega_pore[s]), ("biofilm", omega_biofilm[s])):
            tau = 1.0 / om
            label = "substrate%d in %s: tau = %.6g" % (s, where, tau)
            if tau < TAU_REJECT:
                low.append(label)
            elif tau < TAU_WARN:
                marginal.append(label)
            elif tau > 2.0:
                high.append(label)
    if low:
        raise TauError("a relaxation time below %.2f oscillates instead of diffusing: %s"
                       % (TAU_REJECT, "; ".join(low)))
    return (["above 2 (over-diffusive): " + m for m in high]
            + ["close to 0.5 (may ring): " + m for m in marginal])


def _solve(case, fs: FlowSetup, chunk=0, progress=None):
    """One flow solve to fs.eps or fs.max_iter iterations.  With `chunk` the
    iterations run in calls of at most `chunk` (the convergence tracer carries
    over between calls, so the decisions are those of one call) and
    `progress(iterations so far)` is called after each."""
    if not chunk:
        return case.ns_solve(fs.max_iter, fs.rule, fs.eps, reset=True)
    total, cv = 0, False
    while total < fs.max_iter:
        want = min(chunk, fs.max_iter - total)
        it, cv = case.ns_solve(want, fs.rule, fs.eps, reset=(total == 0))
        total += it
        if progress:
            progress(total)
        if cv or it < want:
            break
    return total, cv


def _after_first_solve(case, fs: FlowSetup):
    """The Peclet rescale after the first solve (complab.cpp:1124-1160): the
    state of the report, and the new pressure drop when a second solve is
    due (None otherwise)."""
    nu = CS2_NS * (fs.ns_tau - 0.5)
    dp = fs.delta_p
    st = case.ns_stats()
    d_lat = CS2_ADE * (fs.tau_ade_fixed - 0.5)
    perm = st[2] * nu * fs.char_length / dp
    u_t = fs.pe * d_lat / fs.char_length
    dp_new = u_t * nu * fs.char_length / perm if perm > 0 else dp
    s = dict(dp=dp, pore_mean=st[3] if fs.has_biofilm else st[0], outlet=st[2], max_ux=st[1],
             perm=perm, d_lat=d_lat)
    return s, (dp_new if abs(dp_new - dp) / dp > 0.01 else None)


def _after_second_solve(case, s, dp_new):
    st = case.ns_stats()
    s.update(pore_mean=st[0], max_ux=st[1], outlet=st[2], dp=dp_new)   # complab.cpp:1175 averages the box


def _flow_report(s, fs: FlowSetup, rep: SetupReport):
    rep.delta_p = s["dp"]
    rep.pore_mean_u = s["pore_mean"]
    rep.max_ux = s["max_ux"]
    rep.outlet_ux = s["outlet"]
    rep.permeability = s["perm"]
    rep.pe_achieved = s["outlet"] * fs.char_length / s["d_lat"]
    return s["pore_mean"]


def solve_flow(case, fs: FlowSetup, rep: SetupReport, chunk=0, progress=None):
    """NS solve + Pe rescale (complab.cpp:1098-1183)."""
    case.ns_init(fs.delta_p)
    it, cv = _solve(case, fs, chunk, progress)
    rep.iters.append(it); rep.converged.append(cv)
    s, dp_new = _after_first_solve(case, fs)
    if dp_new is not None:
        case.ns_init(dp_new)
        it, cv = _solve(case, fs, chunk, progress)
        rep.iters.append(it); rep.converged.append(cv)
        _after_second_solve(case, s, dp_new)
    return _flow_report(s, fs, rep)


def ade_relaxation(p, pore_mean_u, fs: FlowSetup, ts: TransportSetup, rep: SetupReport):
    """Lattice viscosity, relaxation and time step per species (complab.cpp:1220-1252),
    then the tau screen (complab.cpp:1315-1378)."""
    if fs.pe > 1e-12:
        ref_nu = pore_mean_u * fs.char_length / fs.pe
        ref_tau = ref_nu / CS2_ADE + 0.5
        if ref_tau > 2 or ref_tau <= 0.5:
            raise ValueError("ADE tau=%g invalid (complab.cpp:1224)" % ref_tau)
    else:
        ref_tau = ts.tau_if_no_flow
        ref_nu = CS2_ADE * (ref_tau - 0.5)
    if not ts.d_pore[0] > 0:
        raise ValueError("substrate 0 needs a positive pore diffusivity (complab.cpp:1239)")
    dt = ref_nu * ts.dx * ts.dx / ts.d_pore[0]
    omp, omb = [], []
    for s in range(p.n_species):
        nu_s = ref_nu * ts.d_pore[s] / ts.d_pore[0]
        omp.append(1.0 / (nu_s / CS2_ADE + 0.5) if s > 0 else 1.0 / ref_tau)
        omb.append(1.0 / (ref_nu * ts.d_biofilm[s] / ts.d_pore[0] / CS2_ADE + 0.5))
    rep.ref_tau = ref_tau
    rep.dt = dt
    rep.omega_pore = omp
    rep.omega_biofilm = omb
    rep.tau_notes = tau_screen(p, omp, omb)
    return omp, omb, dt


def setup(case, p, fs: FlowSetup, ts: TransportSetup, chunk=0, progress=None):
    """Runs the whole set-up on `case`; returns the report.  `chunk` and
    `progress` split the flow solves as `_solve` does."""
    rep = SetupReport()
    pore_mean = 0.0
    if fs.pe > 0 and p.flow_on:
        pore_mean = solve_flow(case, fs, rep, chunk, progress)
    omp, omb, dt = ade_relaxation(p, pore_mean, fs, ts, rep)
    case.set_omegas(omp, omb, dt)
    if fs.pe > 0 and p.flow_on:
        case.ns_couple()
        if ts.stabilize_steps > 0:
            case.ade_stabilize(ts.stabilize_steps)
    return rep


def setup_batch(ctx, cases, workloads):
    """setup() of many GPU cases (workloads[i] belongs to cases[i]) with their
    flow solves and stabilizations in batch launches (complab_amd.Batch):
    every case ends where setup() would take it.  Returns the reports."""
    from . import Batch, ClbError
    n = len(cases)
    reps = [SetupReport() for _ in range(n)]
    flow = [i for i in range(n) if workloads[i].flow.pe > 0 and workloads[i].params.flow_on]

    def solve(dps):
        for i, dp in dps.items():
            cases[i].ns_init(dp)
        groups = {}
        for i in dps:
            fs = workloads[i].flow
            groups.setdefault((fs.max_iter, fs.rule, fs.eps), []).append(i)
        for (max_iter, rule, eps), ids in groups.items():
            b = Batch(ctx, [cases[i] for i in ids])
            try:
                res = b.ns_solve(max_iter, rule, eps, reset=True)
            finally:
                b.close()
            for i, (it, cv, st) in zip(ids, res):
                if st < 0:
                    raise ClbError(st, "flow solve of %s" % workloads[i].name)
                reps[i].iters.append(it); reps[i].converged.append(cv)

    solve({i: workloads[i].flow.delta_p for i in flow})
    state, again = {}, {}
    for i in flow:
        state[i], dp_new = _after_first_solve(cases[i], workloads[i].flow)
        if dp_new is not None:
            again[i] = dp_new
    if again:
        solve(again)
        for i, dp_new in again.items():
            _after_second_solve(cases[i], state[i], dp_new)
    stab = {}
    for i in range(n):
        w = workloads[i]
        pore_mean = _flow_report(state[i], w.flow, reps[i]) if i in state else 0.0
        omp, omb, dt = ade_relaxation(w.params, pore_mean, w.flow, w.transport, reps[i])
        cases[i].set_omegas(omp, omb, dt)
        if i in state:
            cases[i].ns_couple()
            if w.transport.stabilize_steps > 0:
                stab.setdefault(w.transport.stabilize_steps, []).append(i)
    for steps, ids in stab.items():
        b = Batch(ctx, [cases[i] for i in ids])
        try:
            b.ade_stabilize(steps)
        finally:
            b.close()
    return reps
