"""A CompLaB.xml run end to end over the library: the set-up sequence, the
coupled steps, a snapshot of every field every save_VTK_interval steps and
after the last step, and the diagnostics rows with summary.csv
(complab.cpp:1013-1283 set-up; 2178-2290 the loop's output block, taken at
the top of step iT, i.e. after iT whole steps; 3270-3310 the final output).

    python -m complab_amd.runner path/to/CompLaB.xml [--out DIR] [--steps N]
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


def run(r: inputs.Run, ctx=None, case=None, out_dir: Optional[str] = None, steps: Optional[int] = None,
        log=None) -> Result:
    """Runs `r` on a new device case (or on `case`, created from the same
    parameters and geometry) and writes its output under `out_dir` (default:
    the XML's output path)."""
    t0 = time.perf_counter()
    steps = r.steps if steps is None else steps
    out = out_dir or r.output_path
    if case is None:
        case = Case(ctx or Context(0), r.params, r.geometry)
    say = (lambda m: print(m, file=log, flush=True)) if log else (lambda m: None)
    rep = driver.setup(case, r.params, r.flow, r.transport)
    case.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
    say("set-up done: flow iterations %s, dt %.6g s" % (list(rep.iters), rep.dt))

    flow = bool(r.params.flow_on) and r.flow.pe > 0
    snaps = Snapshots(out, r.species, r.microbes, r.mask_filename, flow=flow)
    diag = Diagnostics(tol=r.diag_tol, enabled=r.diag_enabled)
    for c in r.conserve:
        diag.add_conserve(c)
    csv_path = r.summary_csv if os.path.isabs(r.summary_csv) else os.path.join(out, r.summary_csv)
    csv = SummaryCsv(csv_path) if r.diag_enabled else None
    diag_every = r.diag_interval if r.diag_interval > 0 else max(r.vtk_interval, 0)
    vtk_every = max(r.vtk_interval, 0)

    def due(it, every):
        return every > 0 and it % every == 0

    it = 0
    while True:
        if it < steps:
            if due(it, vtk_every):
                snaps.write(case, it)
            if r.diag_enabled and due(it, diag_every):
                row = row_of(case, it, r.species, r.microbes)
                msg = diag.record(row)
                csv.append(row)
                if msg:
                    say(msg.rstrip("\n"))
        if it >= steps:
            break
        nxt = steps
        for every in (vtk_every, diag_every if r.diag_enabled else 0):
            if every > 0:
                nxt = min(nxt, (it // every + 1) * every)
        rc, n = case.run_steps(nxt - it)
        it += n
        if rc:
            raise RuntimeError("the run stopped with status %d after %d steps" % (rc, it))
        say("step %d" % it)
    snaps.write(case, it)                           # the final output, after the last step
    return Result(steps=it, out_dir=out, files=list(snaps.written), summary_csv=csv_path if csv else None,
                  verdict=diag.verdict(), flow_iters=list(rep.iters), seconds=time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("xml")
    ap.add_argument("--geometry", default=None, help="geometry file (default: the one the XML names)")
    ap.add_argument("--out", default=None, help="output folder (default: the XML's output path)")
    ap.add_argument("--steps", type=int, default=None, help="coupled steps (default: the XML's ade_max_iT)")
    ap.add_argument("--order", default="slices", choices=["slices", "x_fastest"])
    args = ap.parse_args()
    r = inputs.load(args.xml, args.geometry, order=args.order)
    res = run(r, out_dir=args.out, steps=args.steps, log=sys.stderr)
    print("%d steps in %.1f s; %d files in %s; diagnostics %s"
          % (res.steps, res.seconds, len(res.files), res.out_dir, res.verdict or "off"))


if __name__ == "__main__":
    main()
