"""The precision form: one seeded case run in FP64 and in a reduced precision,
compared field by field.

    python -m complab_amd.study --scale 0.25 --steps 5000

Reports, per species, the largest relative difference of the final field over
the open voxels of the reaction box (scaled by the species' largest FP64
magnitude), the final mask difference (voxels whose material differs), and the
drift of each species' total -- |total_low - total_fp64| / |total_fp64| at the
end, scaled to 1e4 steps.  The equilibrium solve, the reductions and the Peclet
rescale are FP64 in every precision.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import Case, Context
from . import configs, driver
from . import params as P


def run(ctx, w, steps, chunk=500, log=None):
    """Set-up and `steps` coupled steps of workload w; with `log`, a line per
    20 000 flow iterations and per `chunk` steps."""
    c = Case(ctx, w.params, w.geometry)
    t0 = time.perf_counter()
    say = (lambda m: print("%s: %s" % (w.name.split()[0], m), file=log, flush=True)) if log else None
    rep = driver.setup(c, w.params, w.flow, w.transport, chunk=20000 if log else 0,
                       progress=(lambda n: say("flow %d iterations" % n)) if log else None)
    c.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
    t1 = time.perf_counter()
    done = 0
    totals = []
    while done < steps:
        rc, n = c.run_steps(min(chunk, steps - done))
        done += n
        totals.append((done, c.species_totals()))
        if say:
            say("%d steps" % done)
        if rc:
            raise RuntimeError("%s stopped with status %d after %d steps" % (w.name, rc, done))
    ctx.sync()
    t2 = time.perf_counter()
    out = {"fields": [c.read(P.FIELD_CONC, s) for s in range(w.params.n_species)],
           "mask": c.read(P.FIELD_MASK), "totals": totals, "counters": c.counters(),
           "setup_s": t1 - t0, "steps_s": t2 - t1, "flow_iters": list(rep.iters)}
    c.close()
    return out


def compare(lo, hi, steps, halo=3):
    """Per species over the open voxels of the reaction box: the largest
    relative difference (scaled by the species' largest FP64 magnitude), its
    50th / 99th / 99.9th percentiles, and the largest one farther than `halo`
    voxels from every voxel whose material differs between the two runs (a
    voxel that sealed or reopened a check apart changes the pore space and
    the flow around it)."""
    from scipy import ndimage
    mask = hi["mask"]
    om = mask >= P.PORE
    om[:, :, 0] = False
    om[:, :, -1] = False
    differ = lo["mask"] != hi["mask"]
    near = ndimage.binary_dilation(differ, iterations=halo) if differ.any() else differ
    far = om & ~near
    per, pct, away = [], [], []
    for a, b in zip(lo["fields"], hi["fields"]):
        scale = max(float(np.max(np.abs(b[om]))), 1e-300)
        e = np.abs(a - b) / scale
        per.append(float(np.max(e[om])))
        pct.append([float("%.3g" % q) for q in np.percentile(e[om], [50, 99, 99.9])])
        away.append(float(np.max(e[far])) if far.any() else 0.0)
    tl, th = lo["totals"][-1][1], hi["totals"][-1][1]
    drift = [float(abs(x - y) / max(abs(y), 1e-300)) * 1e4 / steps for x, y in zip(tl, th)]
    return {"field_max_rel_err": [float("%.3g" % e) for e in per],
            "field_rel_err_p50_p99_p999": pct,
            "field_max_rel_err_beyond_%d_voxels_of_a_mask_change" % halo: [float("%.3g" % e) for e in away],
            "mask_voxels_differ": int(np.sum(differ)),
            "total_drift_per_1e4_steps": [float("%.3g" % d) for d in drift]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=0.25)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--components", type=int, choices=(3, 6), default=3,
                    help="the tableau: 3 components with calcite precipitation and dissolution (the precision "
                         "form), or 6 components on eight aqueous lattices without a mineral")
    ap.add_argument("--precisions", default="fp32_ade,fp32",
                    help="reduced precisions to compare, of fp32_ade and fp32 (empty: the FP64 run alone)")
    ap.add_argument("--flow-max-iter", type=int, default=0,
                    help="cap on every flow solve, the first and the re-solves (0: the workload's own); "
                         "FP64 and FP32-transport runs share that FP64 flow, so a cap keeps the comparison "
                         "while it bounds the run (the full size needs ~100 000 iterations per set-up solve)")
    args = ap.parse_args()

    build = configs.config3 if args.components == 3 else configs.config3_six_components

    def workload(prec):
        w = build(args.scale, prec, args.seed)
        if args.flow_max_iter > 0:
            w.flow.max_iter = args.flow_max_iter
            w.params.ns_max_iter = args.flow_max_iter
        return w

    ctx = Context(0)
    res = {"steps": args.steps, "scale": args.scale, "flow_max_iter": args.flow_max_iter}
    hi = run(ctx, workload(P.FP64), args.steps, log=sys.stderr)
    print("fp64 done: setup %.1f s, steps %.1f s" % (hi["setup_s"], hi["steps_s"]), file=sys.stderr, flush=True)
    res["workload"] = build(args.scale, P.FP64, args.seed).name
    res["fp64"] = {"setup_s": round(hi["setup_s"], 2), "steps_s": round(hi["steps_s"], 2),
                   "flow_iters": hi["flow_iters"], "counters": [int(x) for x in hi["counters"]]}
    chosen = [x for x in args.precisions.split(",") if x]
    for label, prec in (("fp32_ade", P.FP32_ADE), ("fp32", P.FP32)):
        if label not in chosen:
            continue
        lo = run(ctx, workload(prec), args.steps, log=sys.stderr)
        r = compare(lo, hi, args.steps)
        r.update({"setup_s": round(lo["setup_s"], 2), "steps_s": round(lo["steps_s"], 2),
                  "counters": [int(x) for x in lo["counters"]]})
        res[label] = r
        print("%s done: %s" % (label, json.dumps(r)), file=sys.stderr, flush=True)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
