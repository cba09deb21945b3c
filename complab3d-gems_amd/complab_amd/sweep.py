"""Many cases on one GPU: the batched sweep and the serving loop.

A case is independent of every other: its fields stay device-resident from
creation to its last step (flow re-solves included), and its kernels run on
its worker's own HIP stream.  Workers are host threads, one HIP context (one
stream) each; the C calls release Python's GIL, so the device sees the kernels
of every running case interleaved and overlaps them -- many small cases fill
the 256 CUs that one of them leaves mostly idle.

Admission is by HBM footprint (clb_case_footprint, an upper bound): a case
starts when a worker is free and its footprint fits in what the running cases
leave of the budget.  In the serving loop requests arrive on their own clock
and are admitted as workers and memory free up (continuous batching, not a
queue drained in order of a batch); each request's time to solution runs from
its arrival to its last step.

Results do not depend on the mix: every reduction a case makes is a fixed-order
reduction over its own fields, so a case run alone and the same case run
beside others give bit-identical fields.
"""
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import Case, Context, footprint
from . import driver
from . import params as P


@dataclass
class CaseResult:
    name: str
    seconds: float = 0.0            # set-up + steps, on its worker
    setup_seconds: float = 0.0      # of which the set-up (flow solves, stabilization)
    create_seconds: float = 0.0     # of which the case creation (host tables, uploads)
    latency: float = 0.0            # serving: arrival to completion
    steps: int = 0
    flow_iters: List[int] = field(default_factory=list)
    fields: Optional[List[np.ndarray]] = None   # final mobile species fields (float64)
    open_mask: Optional[np.ndarray] = None
    status: int = 0
    error: str = ""


def run_workload(ctx, w, steady_tol=None, check_every=100, keep_fields=True):
    """Creates the case, runs the set-up (flow solve, Peclet rescale, ADE
    relaxation, stabilization) and its steps -- w.steps of them, or until the
    mobile species totals change by less than steady_tol (relative) between two
    checks check_every steps apart -- then reads the final fields back."""
    r = CaseResult(w.name)
    t0 = time.perf_counter()
    c = Case(ctx, w.params, w.geometry)
    try:
        rep = driver.setup(c, w.params, w.flow, w.transport)
        r.flow_iters = list(rep.iters)
        c.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
        ctx.sync()
        r.setup_seconds = time.perf_counter() - t0
        if steady_tol is None:
            rc, n = c.run_steps(w.steps)
            r.steps, r.status = n, rc
        else:
            prev = c.species_totals()
            done = 0
            while done < w.steps:
                k = min(check_every, w.steps - done)
                rc, n = c.run_steps(k)
                done += n
                if rc:
                    r.status = rc
                    break
                tot = c.species_totals()
                mob = [s for s in range(w.params.n_species) if w.params.mobile[s]]
                scale = np.maximum(np.abs(tot[mob]), 1e-300)
                if np.max(np.abs(tot[mob] - prev[mob]) / scale) < steady_tol:
                    break
                prev = tot
            r.steps = done
        if keep_fields:
            r.fields = [c.read(P.FIELD_CONC, s) for s in range(w.params.n_species) if w.params.mobile[s]]
            m = c.read(P.FIELD_MASK)
            om = m >= P.PORE
            om[:, :, 0] = False
            om[:, :, -1] = False
            r.open_mask = om
        ctx.sync()
    finally:
        c.close()
    r.seconds = time.perf_counter() - t0
    return r


class _Admission:
    """HBM budget shared by the workers."""

    def __init__(self, budget):
        self.budget = budget
        self.used = 0
        self.cv = threading.Condition()

    def acquire(self, nbytes):
        with self.cv:
            while self.used > 0 and self.used + nbytes > self.budget:
                self.cv.wait()
            self.used += nbytes

    def release(self, nbytes):
        with self.cv:
            self.used -= nbytes
            self.cv.notify_all()


def _pool(jobs, workers, device, budget, steady_tol, keep_fields, arrivals=None, log=None):
    """Runs (index, workload) jobs on `workers` threads; with `arrivals` each
    job enters the queue at its arrival time (seconds after the start).  `log`,
    when given, is called with each finished result."""
    q = queue.Queue()
    results = [None] * len(jobs)
    adm = _Admission(budget)
    start = time.perf_counter()
    err = []

    def worker():
        try:
            ctx = Context(device)
        except Exception as e:          # noqa: BLE001 - reported to the caller
            err.append(e)
            return
        try:
            while True:
                item = q.get()
                if item is None:
                    return
                i, w, t_arr = item
                fp = footprint(w.params)
                adm.acquire(fp)
                try:
                    r = run_workload(ctx, w, steady_tol=steady_tol, keep_fields=keep_fields)
                except Exception as e:  # noqa: BLE001 - reported per case
                    r = CaseResult(w.name, status=-1, error=str(e))
                finally:
                    adm.release(fp)
                r.latency = time.perf_counter() - start - t_arr
                results[i] = r
                if log is not None:
                    log(r)
        finally:
            ctx.close()

    threads = [threading.Thread(target=worker, daemon=True) for _ in range(workers)]
    for t in threads:
        t.start()
    for k, (i, w) in enumerate(jobs):
        t_arr = 0.0
        if arrivals is not None:
            t_arr = arrivals[k]
            delay = start + t_arr - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        q.put((i, w, t_arr))
    for _ in threads:
        q.put(None)
    for t in threads:
        t.join()
    wall = time.perf_counter() - start
    if err:
        raise err[0]
    return results, wall


def run_sequential(workloads, device=0, steady_tol=None, keep_fields=True, log=None):
    """The plain per-case path: one case at a time on one stream."""
    return _pool(list(enumerate(workloads)), 1, device, float("inf"), steady_tol, keep_fields, log=log)


def run_batched(workloads, workers=8, device=0, budget=200e9, steady_tol=None, keep_fields=True):
    """All cases admitted at once, `workers` of them on the device together."""
    return _pool(list(enumerate(workloads)), workers, device, budget, steady_tol, keep_fields)


def serve(workloads, rate_per_s=2.0, workers=8, device=0, budget=200e9, steady_tol=1e-6, seed=0,
          keep_fields=True):
    """The serving loop: requests arrive as a Poisson stream of `rate_per_s`
    and are admitted as workers and HBM free up.  Returns the results (with
    each request's arrival-to-completion latency) and the wall time."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_per_s, size=len(workloads)))
    arrivals -= arrivals[0]
    return _pool(list(enumerate(workloads)), workers, device, budget, steady_tol, keep_fields, arrivals=list(arrivals))


def _results_of(c, w, r, keep_fields):
    if keep_fields:
        r.fields = [c.read(P.FIELD_CONC, s) for s in range(w.params.n_species) if w.params.mobile[s]]
        m = c.read(P.FIELD_MASK)
        om = m >= P.PORE
        om[:, :, 0] = False
        om[:, :, -1] = False
        r.open_mask = om


def run_group(workloads, device=0, keep_fields=True, create_threads=8, log=None):
    """The batched form: the cases are created together (their host tables
    built in parallel threads), their set-up solves and stabilizations run in
    batch launches (driver.setup_batch), then their steps (clb_batch_run_steps,
    each case's own `steps`).  Returns the results and the wall time; a
    result's `seconds` is the wall time at which its case finished."""
    from concurrent.futures import ThreadPoolExecutor
    from . import Batch
    start = time.perf_counter()
    ctx = Context(device)
    results = [CaseResult(w.name) for w in workloads]
    cases = []
    try:
        with ThreadPoolExecutor(max(1, create_threads)) as ex:
            cases = list(ex.map(lambda w: Case(ctx, w.params, w.geometry), workloads))
        ctx.sync()
        t_create = time.perf_counter() - start
        reps = driver.setup_batch(ctx, cases, workloads)
        for c, rep, r in zip(cases, reps, results):
            c.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
            r.flow_iters = list(rep.iters)
        ctx.sync()
        t_setup = time.perf_counter() - start
        b = Batch(ctx, cases)
        live = list(range(len(cases)))
        try:
            while live:
                k = min(workloads[i].steps - results[i].steps for i in live)
                out = b.run_steps(k)
                still = []
                for i, (rc, d) in zip(live, out):
                    r = results[i]
                    r.steps += d
                    if rc:
                        r.status = rc
                    if rc or r.steps >= workloads[i].steps:
                        b.remove(cases[i])
                        _results_of(cases[i], workloads[i], r, keep_fields)
                        r.setup_seconds = t_setup
                        r.create_seconds = t_create
                        r.seconds = r.latency = time.perf_counter() - start
                        if log is not None:
                            log(r)
                    else:
                        still.append(i)
                live = still
        finally:
            b.close()
        ctx.sync()
    finally:
        for c in cases:
            c.close()
        ctx.close()
    return results, time.perf_counter() - start


def _steady(tot, prev, mobile, tol):
    scale = np.maximum(np.abs(tot[mobile]), 1e-300)
    return bool(np.max(np.abs(tot[mobile] - prev[mobile]) / scale) < tol)


def serve_group(workloads, rate_per_s=2.0, device=0, budget=200e9, steady_tol=1e-6, check_every=100, seed=0,
                keep_fields=True, create_threads=8, log=None):
    """The serving loop over the batched form (continuous batching): requests
    arrive as a Poisson stream of `rate_per_s`; between runs of `check_every`
    steps the requests that have arrived and whose footprint fits in what the
    running ones leave of `budget` are created, set up together
    (driver.setup_batch) and join the running batch, and members that reached
    steady state (mobile totals changing by less than `steady_tol`, relative,
    over the run), their step limit or a stop leave it.  Every member checks at
    multiples of `check_every` of its own steps, as run_workload does, so each
    result equals its one-at-a-time run.  Returns the results (latency =
    arrival to last step) and the wall time."""
    from collections import deque
    from concurrent.futures import ThreadPoolExecutor
    from . import Batch
    n = len(workloads)
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))
    arrivals -= arrivals[0]
    results = [CaseResult(w.name) for w in workloads]
    pending = deque(range(n))
    running = {}                     # request -> [case, totals at the last check, footprint]
    used = 0
    start = time.perf_counter()
    ctx = Context(device)
    b = Batch(ctx, precision=workloads[0].params.precision) if n else None
    try:
        with ThreadPoolExecutor(max(1, create_threads)) as ex:
            while pending or running:
                now = time.perf_counter() - start
                new = []
                while pending and arrivals[pending[0]] <= now:
                    fp = footprint(workloads[pending[0]].params)
                    if (running or new) and used + fp > budget:
                        break
                    used += fp
                    new.append((pending.popleft(), fp))
                if not running and not new:
                    time.sleep(max(0.0, arrivals[pending[0]] - now))
                    continue
                if new:
                    ids = [i for i, _ in new]
                    cases = list(ex.map(lambda i: Case(ctx, workloads[i].params, workloads[i].geometry), ids))
                    reps = driver.setup_batch(ctx, cases, [workloads[i] for i in ids])
                    for (i, fp), c, rep in zip(new, cases, reps):
                        c.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
                        results[i].flow_iters = list(rep.iters)
                        results[i].setup_seconds = time.perf_counter() - start - arrivals[i]
                        b.add(c)
                        running[i] = [c, c.species_totals(), fp]
                order = [next(i for i, v in running.items() if v[0] is c) for c in b.cases]

                def next_check(i):
                    s = results[i].steps
                    return min((s // check_every + 1) * check_every, workloads[i].steps)
                k = min(next_check(i) - results[i].steps for i in order)
                due = {i: next_check(i) for i in order}
                out = b.run_steps(k)
                for i, (rc, d) in zip(order, out):
                    c, prev, fp = running[i]
                    r, w = results[i], workloads[i]
                    r.steps += d
                    done = rc != 0 or r.steps >= w.steps
                    if rc:
                        r.status = rc
                    if not done and r.steps == due[i]:
                        tot = c.species_totals()
                        mob = [s for s in range(w.params.n_species) if w.params.mobile[s]]
                        done = _steady(tot, prev, mob, steady_tol)
                        running[i][1] = tot
                    if done:
                        b.remove(c)
                        _results_of(c, w, r, keep_fields)
                        c.close()
                        del running[i]
                        used -= fp
                        r.latency = time.perf_counter() - start - arrivals[i]
                        r.seconds = r.latency
                        if log is not None:
                            log(r)
    finally:
        for v in running.values():
            try:
                b.remove(v[0])
            except Exception:       # noqa: BLE001 - tearing down after an error
                pass
            v[0].close()
        if b is not None:
            b.close()
        ctx.close()
    return results, time.perf_counter() - start


def field_error(a, b):
    """Largest relative difference of the final mobile fields of two runs of a
    case, over the open voxels of the reaction box, scaled per species by the
    largest magnitude of `b`."""
    e = 0.0
    for fa, fb in zip(a.fields, b.fields):
        m = b.open_mask
        scale = max(float(np.max(np.abs(fb[m]))), 1e-300)
        e = max(e, float(np.max(np.abs(fa[m] - fb[m]))) / scale)
    return e


def percentile(values, q):
    return float(np.percentile(np.asarray(values, dtype=float), q)) if len(values) else 0.0
