"""A CompLaB.xml run end to end over the library: the set-up sequence, the
coupled steps, a snapshot of every field every save_VTK_interval steps and
after the last step, the diagnostics rows with summary.csv, and a checkpoint
every save_CHK_interval steps (complab.cpp:1013-1283 set-up; 2178-2302 the
loop's output block, taken at the top of step iT, i.e. after iT whole steps;
3270-3310 the final output).  A run restarted from a checkpoint skips the
set-up and continues from the step the checkpoint holds, bit-identically.

    python -m complab_amd.runner path/to/CompLaB.xml [--out DIR] [--steps N] [--restart FILE]
"""
import argparse
import os
import sys
import time
from dataclasses import dataclass, field
from typing import List, Optional

from . import Case, Context
from . import driver
from . import inputs
from .diagnostics import Diagnostics, row_of
from .output import Snapshots, SummaryCsv


@dataclass
class Result:
    steps: int
    out_dir: str
    files: List[str] = field(default_factory=list)
    summary_csv: Optional[str] = None
    verdict: str = ""
    flow_iters: List[int] = field(default_factory=list)
    seconds: float = 0.0
    checkpoints: List[str] = field(default_factory=list)


CHK_STEP = 7                                    # counters(): steps run


def checkpoint_name(iteration: int) -> str:
    return "checkpoint_%07d.clbchk" % iteration


def vtk_every_of(r) -> int:
    return max(r.vtk_interval, 0)



def gated_microbes(r):
    """The organisms that have a reaction block of the .thm file, as
    (index, name): one pair of dG_/FT_ snapshot files each
    (gatedMicrobes, complab3d_energyfield.hh:190-197)."""
    pr = r.params
    if not pr.thm_on:
        return []
    return [(m, r.microbes[m]) for m in range(len(r.microbes)) if pr.thm_microbe < 0 or m == pr.thm_microbe]


def run(r: inputs.Run, ctx=None, case=None, out_dir: Optional[str] = None, steps: Optional[int] = None,
        log=None, restart: Optional[str] = None) -> Result:
    """Runs `r` on a new device case (or on `case`, created from the same
    parameters and geometry) and writes its output under `out_dir` (default:
    the XML's output path).  With `restart`, the case is loaded from that
    checkpoint file instead of being set up."""
    t0 = time.perf_counter()
    steps = r.steps if steps is None else steps
    out = out_dir or r.output_path
    if case is None:
        case = Case(ctx or Context(0), r.params, r.geometry)
    for m, prog in r.symbolic.items():
        case.set_symbolic(m, prog)
    say = (lambda m: print(m, file=log, flush=True)) if log else (lambda m: None)
    iters = []
    if restart:
        case.load_checkpoint(restart)
        say("restarted from %s at step %d" % (restart, int(case.counters()[CHK_STEP])))
    else:
        rep = driver.setup(case, r.params, r.flow, r.transport)
        case.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
        iters = list(rep.iters)
        say("set-up done: flow iterations %s, dt %.6g s" % (iters, rep.dt))

    flow = bool(r.params.flow_on) and r.flow.pe > 0
    pr = r.params
    gated = gated_microbes(r) if vtk_every_of(r) > 0 else []
    snaps = Snapshots(out, r.species, r.microbes, r.mask_filename, flow=flow, energy=gated)
    diag = Diagnostics(tol=r.diag_tol, enabled=r.diag_enabled)
    for c in r.conserve:
        diag.add_conserve(c)
    csv_path = r.summary_csv if os.path.isabs(r.summary_csv) else os.path.join(out, r.summary_csv)
    csv = SummaryCsv(csv_path) if r.diag_enabled else None
    diag_every = r.diag_interval if r.diag_interval > 0 else max(r.vtk_interval, 0)
    vtk_every = max(r.vtk_interval, 0)
    chk_every = max(r.chk_interval, 0)
    checkpoints = []

    def due(it, every):
        return every > 0 and it % every == 0

    # rate snapshots (complab.cpp:2078-2105, 2548-2561, 3284): with a rate law
    # on, the step after each concentration snapshot is captured and written
    # as rate_<species>_<iT>.vti under the snapshot's number, and the last
    # step under the final snapshot's
    pr = r.params
    has_law = (pr.bio_law != 0 and pr.n_microbes > 0) or pr.n_rxn > 0 or len(r.symbolic) > 0
    rates = vtk_every > 0 and not pr.dissol_on and has_law
    if vtk_every > 0 and pr.dissol_on and has_law:
        say("rate snapshots are not written for a run with dissolution: the rate capture "
            "does not cover the dissolution kinetics on this path")
    capture_at = -1

    it = int(case.counters()[CHK_STEP]) if restart else 0
    first = it
    while True:
        if it < steps:
            if due(it, vtk_every):
                snaps.write(case, it)
                if rates:
                    capture_at = it
            if r.diag_enabled and due(it, diag_every):
                row = row_of(case, it, r.species, r.microbes)
                msg = diag.record(row)
                csv.append(row)
                if msg:
                    say(msg.rstrip("\n"))
            if due(it, chk_every) and it > first:          # complab.cpp:2293 (iT > 0)
                path = os.path.join(out, checkpoint_name(it))
                case.save_checkpoint(path)
                checkpoints.append(path)
        if it >= steps:
            break
        nxt = steps
        for every in (vtk_every, diag_every if r.diag_enabled else 0, chk_every):
            if every > 0:
                nxt = min(nxt, (it // every + 1) * every)
        last = rates and it == steps - 1
        capture = capture_at == it or last
        if capture:
            nxt = it + 1
            case.rate_capture(True)
        elif rates and nxt == steps:
            nxt = steps - 1                           # the last step runs on its own
        rc, n = case.run_steps(nxt - it)
        if capture:
            if n == 1 and capture_at == it:
                snaps.write_rates(case, it)
            if not (n == 1 and last):
                case.rate_capture(False)
            capture_at = -1
        it += n
        if rc:
            raise RuntimeError("the run stopped with status %d after %d steps" % (rc, it))
        say("step %d" % it)
    snaps.write(case, it)                           # the final output, after the last step
    if rates and it == steps and it > first:
        snaps.write_rates(case, it)
        case.rate_capture(False)
    return Result(steps=it, out_dir=out, files=list(snaps.written), summary_csv=csv_path if csv else None,
                  verdict=diag.verdict(), flow_iters=iters, seconds=time.perf_counter() - t0,
                  checkpoints=checkpoints)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("xml")
    ap.add_argument("--geometry", default=None, help="geometry file (default: the one the XML names)")
    ap.add_argument("--out", default=None, help="output folder (default: the XML's output path)")
    ap.add_argument("--steps", type=int, default=None, help="coupled steps (default: the XML's ade_max_iT)")
    ap.add_argument("--order", default="slices", choices=["slices", "x_fastest"])
    ap.add_argument("--restart", default=None, help="continue from this checkpoint file")
    args = ap.parse_args()
    r = inputs.load(args.xml, args.geometry, order=args.order)
    res = run(r, out_dir=args.out, steps=args.steps, log=sys.stderr, restart=args.restart)
    print("%d steps in %.1f s; %d files in %s; diagnostics %s"
          % (res.steps, res.seconds, len(res.files), res.out_dir, res.verdict or "off"))


if __name__ == "__main__":
    main()
