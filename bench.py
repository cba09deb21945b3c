"""Benchmark of the coupled reaction step on one MI355X.

Workload: the biotic plain-method configuration -- 256x128x128 voxels (+2 ghost
x columns) seeded sphere pack at porosity 0.35 (r = 6), four D3Q7 species,
dual-Monod biofilm on 2 % of the grain surface, CA spill, flow re-solved on
mask changes every 50 steps, FP64.  The set-up (flow solve, Peclet rescale,
stabilization) is untimed; W warm-up steps, then K coupled steps are timed
between two device synchronisations.  Prints one JSON line.

value: coupled MLUPS = lattice voxels x K / seconds / 1e6.
roofline: the kernel with the most device time in the timed window -- the
flow kernel of the in-step re-solves (19 x 2 x p bytes per open voxel, plus
1 byte of mask per voxel) or the fused transport + kinetics kernel (per open
voxel S_m*14*p + 3p + (S_i+M)*2p, plus 1 byte of mask per voxel) --
algorithmic bytes per launch over its HIP-event-timed average duration,
against 8 TB/s; `traffic` is the HBM bytes per launch of the committed
rocprofv3 counter summary in profiles/.  Both kernels' figures are in
roofline_flow and roofline_transport.
cpu_baseline: the CPU oracle (a scalar C++ restatement) on the same case for a
few steps, velocity copied from the GPU flow solve.
sweep: the batched form -- seeded sweep cases of the batched workload (cuboids
with edges drawn in 32..160, aspect <= 4, spheres / cylinders / fracture,
2-6 species, Monod growth), each through its full set-up and `--sweep-steps`
steps: cases/hour of the plain path (FP64, one case at a time) and of the
batched path (several cases on the device together, FP32 transport with FP64
flow, and all-FP32), with the largest relative field error against the plain
path.
serving: ABC-style requests (A + B -> mineral, precipitation re-solving the
flow every 100 steps) arriving as a Poisson stream at 2/s, admitted by HBM
footprint as workers free up, run to steady state or `--serve-steps`; p50/p99
time to solution (arrival to last step), cases/hour, and the plain path (the
same requests one at a time) with the error between the two.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "complab3d-gems_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# the timed region brackets every 8th transport launch with HIP events: each
# event pair puts markers on the stream that cost the device a dispatch gap
TIMING_STRIDE = 8
PROFILES = {"k_ade_bulk": os.path.join(ROOT, "profiles", "k_ade_bulk_biotic_258x128x128_fp64.txt"),
            "k_ns_step": os.path.join(ROOT, "profiles", "k_ns_step_biotic_258x128x128_fp64.txt")}


def profiled_traffic(p, n_open, kernel):
    """HBM bytes per launch of `kernel`, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated against a copy
    kernel of known size); None for any other workload."""
    if p.precision != 0 or (p.nx, p.ny, p.nz) != (258, 128, 128) or n_open != 1485001:
        return None
    try:
        for line in open(PROFILES[kernel]):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        return None
    return None


def algorithmic_bytes(p, n_open, n_vox):
    """Bytes one transport launch must move: per open voxel 7 populations read
    and 7 written per mobile species, the velocity, each immobile species and
    each biomass read and written, and -- with a CA microbe, whose mask and
    total update the launch makes -- the biomass total read and written; plus
    the 1-byte mask of every voxel."""
    r = 8 if p.precision == 0 else 4
    sm = sum(1 for s in range(p.n_species) if p.mobile[s])
    si = p.n_species - sm
    per_open = sm * 14 * r + 3 * r + (si + p.n_microbes) * 2 * r
    if p.n_microbes > 0:
        per_open += 2 * r
    return per_open * n_open + n_vox, per_open


def cpu_baseline(w, g, rep, steps):
    """The oracle on the same case, velocity copied from the GPU flow solve:
    `steps` coupled steps on every host thread the process has, and one step
    on one thread.  Returns (MLUPS all threads, threads, steps, seconds,
    MLUPS one thread)."""
    import oracle as O
    from complab_amd import params as P
    o = O.OracleCase(w.params, w.geometry)
    for a in range(3):
        o.write(P.FIELD_VEL, a, g.read(P.FIELD_VEL, a))
    o.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
    threads = O.set_threads(0)
    o.run_steps(1)
    V = w.params.nx * w.params.ny * w.params.nz
    t0 = time.perf_counter()
    rc, n = o.run_steps(steps)
    dt = time.perf_counter() - t0
    O.set_threads(1)
    t1 = time.perf_counter()
    o.run_steps(1)
    dt1 = time.perf_counter() - t1
    O.set_threads(threads)
    o.close()
    return V * n / dt / 1e6, threads, n, dt, V / dt1 / 1e6


def abc_one_thread(steps=20):
    """The oracle on one thread on an ABC-style transport step -- 32^3 voxels
    (sphere pack with y/z walls), 3 D3Q7 species, A + B -> C by second-order
    mass action -- for the reference's published 0.050 s per step of that
    case on one core.  Seconds per coupled step."""
    import oracle as O
    from complab_amd import configs
    from complab_amd import params as P
    lat, _ = configs._pack(32, 32, 32, 3.0, 0.36, 1036)
    p = P.default_params(34, 32, 32, 3)
    for s in range(3):
        p.left_btype[s] = P.BC_DIRICHLET; p.right_btype[s] = P.BC_NEUMANN
    p.left_bc[0] = 1.0; p.c0[1] = 1.0
    p.n_rxn = 1; p.rxn_r1[0] = 0; p.rxn_r2[0] = 1; p.rxn_k[0] = 1e-2
    p.rxn_stoich[0][0] = -1.0; p.rxn_stoich[0][1] = -1.0; p.rxn_stoich[0][2] = 1.0
    threads = O.set_threads(0)
    O.set_threads(1)
    try:
        o = O.OracleCase(p, lat)
        o.set_omegas([1.2] * 3, [1.2] * 3, 1.0)
        o.run_steps(2)
        t0 = time.perf_counter()
        o.run_steps(steps)
        dt = (time.perf_counter() - t0) / steps
        o.close()
    finally:
        O.set_threads(threads)
    return dt


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _progress(tag):
    def log(r):
        print("%s: %s done, %d steps, %.2f s" % (tag, r.name, r.steps, r.seconds), file=sys.stderr, flush=True)
    return log


def sweep_measurement(args):
    """The batched form against the plain path: the same seeded cases, each
    through its full set-up and `--sweep-steps` coupled steps, one case at a
    time in FP64 (the plain per-case path) and all together through the
    batched entry points (sweep.run_group: one launch per flow iteration and
    per transport step over every member, graph-replayed runs, the autotuned
    transport workgroup) in FP64, FP32 transport with an FP64 flow, and FP32."""
    from complab_amd import configs, sweep
    from complab_amd import params as P
    n, steps = args.sweep_cases, args.sweep_steps
    plain = [configs.sweep_case(s, P.FP64, steps=steps) for s in range(n)]
    r64, t64 = sweep.run_sequential(plain, log=_progress("sweep plain fp64"))
    if any(r.status for r in r64):
        raise SystemExit("a plain sweep case failed: %s" % [r.error or r.status for r in r64])
    cph64 = n / t64 * 3600.0
    vox = sum(w.params.nx * w.params.ny * w.params.nz for w in plain)
    res = {"cases": n, "steps": steps, "voxels_total": int(vox),
           "workload": "seeded batched-form cases: edges 32..160, aspect <= 4, spheres/cylinders/fracture, "
                       "porosity 0.25-0.5, 2-6 species, flow tau 0.6-1.2, Monod growth; full set-up per case",
           "plain_fp64": {"wall_s": round(t64, 3), "cases_per_hour": round(cph64, 1)}}
    for label, prec in (("fp64", P.FP64), ("fp32_ade", P.FP32_ADE), ("fp32", P.FP32)):
        ws = [configs.sweep_case(s, prec, steps=steps) for s in range(n)]
        r, t = sweep.run_group(ws, log=_progress("sweep batched " + label))
        if any(x.status for x in r):
            raise SystemExit("a batched sweep case failed: %s" % [x.error or x.status for x in r])
        err = max(sweep.field_error(a, b) for a, b in zip(r, r64))
        res["batched_" + label] = {"wall_s": round(t, 3), "cases_per_hour": round(n / t * 3600.0, 1),
                                   "speedup_vs_plain_fp64": round(t64 / t, 3), "max_rel_err": float("%.3g" % err),
                                   "setup_s": round(max(x.setup_seconds for x in r), 3),
                                   "create_s": round(max(x.create_seconds for x in r), 3)}
        print("sweep: batched %s %.1f s (plain %.1f s), error %.3g" % (label, t, t64, err), file=sys.stderr, flush=True)
    return res


def serving_measurement(args):
    """Continuous batching (sweep.serve_group) of ABC-style requests in FP32
    transport with an FP64 flow, against the plain path: the same requests in
    FP64, one at a time.  Precipitation seals a voxel at a threshold, so a
    reduced precision can seal it a check earlier or later and change the pore
    space: the line reports the error of every request and how many are within
    1e-5."""
    from complab_amd import configs, sweep
    from complab_amd import params as P
    n = args.serve_requests
    size = (48, max(48, args.serve_size))
    reqs = [configs.serving_case(100 + s, P.FP32_ADE, size=size, max_steps=args.serve_steps) for s in range(n)]
    r, wall = sweep.serve_group(reqs, rate_per_s=args.serve_rate, steady_tol=1e-6, seed=7,
                                log=_progress("serving"))
    if any(x.status not in (0, 1) for x in r):
        raise SystemExit("a serving request failed: %s" % [x.error or x.status for x in r])
    lat = [x.latency for x in r]
    plain_w = [configs.serving_case(100 + s, P.FP64, size=size, max_steps=args.serve_steps) for s in range(n)]
    plain, t_plain = sweep.run_sequential(plain_w, steady_tol=1e-6, log=_progress("serving plain fp64"))
    errs = [sweep.field_error(a, b) for a, b in zip(r, plain)]
    print("serving: wall %.1f s, p50 %.2f s, p99 %.2f s, plain one-at-a-time %.1f s, worst error %.3g" %
          (wall, sweep.percentile(lat, 50), sweep.percentile(lat, 99), t_plain, max(errs)), file=sys.stderr, flush=True)
    return {"requests": n, "rate_per_s": args.serve_rate, "precision": "fp32 transport, fp64 flow",
            "workload": "ABC-style A+B->mineral, precipitation with flow re-solve every 100 steps, edges %d..%d, "
                        "3-5 species, to steady state (1e-6) or %d steps" % (size[0], size[1], args.serve_steps),
            "edges": [[w.params.nx - 2, w.params.ny, w.params.nz] for w in reqs],
            "p50_s": round(sweep.percentile(lat, 50), 3), "p99_s": round(sweep.percentile(lat, 99), 3),
            "wall_s": round(wall, 3), "cases_per_hour": round(n / wall * 3600.0, 1),
            "plain_fp64_wall_s": round(t_plain, 3), "plain_fp64_cases_per_hour": round(n / t_plain * 3600.0, 1),
            "speedup_vs_plain_fp64": round(t_plain / wall, 3),
            "steps": [x.steps for x in r], "plain_steps": [x.steps for x in plain],
            "rel_err_vs_plain_fp64": [float("%.3g" % e) for e in errs],
            "within_1e-5": int(sum(e <= 1e-5 for e in errs))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stabilize", type=int, default=10000)
    ap.add_argument("--ns-max-iter", type=int, default=100000,
                    help="cap of the set-up flow solves (profiling runs shorten the untimed set-up)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched and serving measurements")
    ap.add_argument("--sweep-cases", type=int, default=64)
    ap.add_argument("--sweep-steps", type=int, default=1000)
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-steps", type=int, default=5000)
    ap.add_argument("--serve-size", type=int, default=256, help="largest edge of a serving request")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="request arrivals per second")
    ap.add_argument("--max-to-spill", type=int, default=3000,
                    help="untimed steps allowed after the warm-up to reach the CA spill")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(each case is one device's work; see the sharding design notes)", file=sys.stderr)
        sys.exit(2)

    import complab_amd as clb
    from complab_amd import configs, driver
    from complab_amd import params as P

    prec = P.FP32 if args.precision == "fp32" else P.FP64
    w = configs.config2(scale=args.scale, precision=prec)
    w.transport.stabilize_steps = args.stabilize
    w.flow.max_iter = args.ns_max_iter
    ctx = clb.Context(0)
    dev = ctx.device_name()
    g = clb.Case(ctx, w.params, w.geometry)
    t0 = time.perf_counter()
    rep = driver.setup(g, w.params, w.flow, w.transport)
    g.set_omegas(rep.omega_pore, rep.omega_biofilm, 1.0)
    setup_s = time.perf_counter() - t0
    print("setup: %.1f s, flow iterations %s, dP %.4g, Pe %.3f, dt %.4g" %
          (setup_s, rep.iters, rep.delta_p, rep.pe_achieved, rep.dt), file=sys.stderr, flush=True)

    rc, n = g.run_steps(args.warmup)
    if rc:
        raise SystemExit("warm-up stopped with status %d after %d steps" % (rc, n))
    # untimed: advance until the CA spill is under way, so that any timed
    # window holds the whole coupled step (spill sweeps, mask changes) and not
    # only the transport of the biofilm's first growth
    to_spill = 0
    while g.counters()[6] == 0 and to_spill < args.max_to_spill:
        rc, n = g.run_steps(100)
        if rc:
            raise SystemExit("advance to the spill stopped with status %d" % rc)
        to_spill += n
    print("spill under way after %d steps past the warm-up" % to_spill, file=sys.stderr, flush=True)
    ctx.sync()
    g.set_timing_stride(TIMING_STRIDE)
    g.kernel_time(reset=True)
    g.flow_kernel_time(reset=True)
    c0 = g.counters()
    # a small copy-kernel launch on each side of the timed region marks it in
    # kernel traces (k_copy16 runs nowhere else before the copy measurement)
    ctx.stream_copy(1 << 20, 1)
    ctx.sync()
    t0 = time.perf_counter()
    rc, n = g.run_steps(args.steps)
    ctx.sync()
    t = time.perf_counter() - t0
    ctx.stream_copy(1 << 20, 1)
    if rc or n != args.steps:
        raise SystemExit("timed run stopped with status %d after %d steps" % (rc, n))
    c1 = g.counters()
    kms, launches = g.kernel_time()
    fms, flaunches = g.flow_kernel_time()

    # the flow: a re-solve as the main loop runs one after a mask change
    # (its own tracer starts empty: complab.cpp:2166), timed alone, and its
    # cost per coupled step at the configured re-solve interval; then the
    # per-iteration time of the flow kernel and tracer
    ctx.sync()
    t1 = time.perf_counter()
    g.ns_update_dynamics(w.params.perm_ratio)
    rs_it, rs_cv = g.ns_solve(w.params.ns_max_iter, w.params.ns_conv_rule, w.params.ns_conv_eps, reset=True)
    ctx.sync()
    rs_s = time.perf_counter() - t1
    t1 = time.perf_counter()
    fl_it, _ = g.ns_solve(500, P.CONV_TRACER, 1e-300, reset=True)      # a tolerance no window meets
    ctx.sync()
    fl_us = (time.perf_counter() - t1) / max(fl_it, 1) * 1e6
    copy_gbs, copy_ms = ctx.stream_copy(1 << 30, 20)
    p = w.params
    V = p.nx * p.ny * p.nz
    mask = g.read(P.FIELD_MASK)
    n_open = int((mask >= 2).sum())
    # the timed kernel works on the interior voxels x = 1..nx-2
    n_int = int((mask[:, :, 1:-1] >= 2).sum())
    bytes_launch, per_open = algorithmic_bytes(p, n_int, (p.nx - 2) * p.ny * p.nz)
    achieved = bytes_launch / (kms * 1e-3) / 1e9
    mlups = V * args.steps / t / 1e6
    ns_bytes = 19 * 2 * (8 if p.precision != P.FP32 else 4) * n_open + V    # per flow iteration
    transport = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(p, n_open, "k_ade_bulk"),
                 "kernel": "k_ade_bulk (fused stream+collide+kinetics of the interior voxels)",
                 "kernel_avg_ms": round(kms, 5), "launches": int(launches), "timed_every": TIMING_STRIDE,
                 "bytes_per_open_voxel": per_open, "bytes_per_launch": int(bytes_launch),
                 "device_ms_in_window": round(kms * args.steps, 2)}
    # the flow kernel: working iterations of the in-step re-solves, every
    # TIMING_STRIDE-th timed; its device time in the window is the timed mean
    # times the working iterations (timed count x stride)
    flow = None
    if flaunches > 0:
        f_ach = ns_bytes / (fms * 1e-3) / 1e9
        flow = {"bound": "hbm", "achieved": round(f_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(f_ach / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(p, n_open, "k_ns_step"),
                "kernel": "k_ns_step (one D3Q19 iteration: pull, incompressible BGK, bounce-back, pressure planes)",
                "kernel_avg_ms": round(fms, 5), "launches": int(flaunches), "timed_every": TIMING_STRIDE,
                "bytes_per_open_voxel": 19 * 2 * (8 if p.precision != P.FP32 else 4),
                "bytes_per_launch": int(ns_bytes),
                "device_ms_in_window": round(fms * flaunches * TIMING_STRIDE, 2)}
    dominant = transport
    if flow is not None and flow["device_ms_in_window"] > transport["device_ms_in_window"]:
        dominant = flow
    resolve_ms_per_step = rs_s * 1e3 / max(p.ns_update_interval, 1)

    cpu = None
    if not args.no_cpu:
        cv, threads, cn, cdt, cv1 = cpu_baseline(w, g, rep, args.cpu_steps)
        abc = abc_one_thread()
        cpu = {"value": round(cv, 4), "unit": "MLUPS", "cores": threads, "kind": "port",
               "value_one_thread": round(cv1, 4), "cpu_model": cpu_model(),
               "abc_32cube_3species_s_per_step_one_thread": round(abc, 5),
               "reference_published_abc_s_per_step_one_core": 0.050,
               "sample": "%d coupled steps of the same %dx%dx%d case by the C++ oracle (OpenMP over voxels, "
                         "%d threads; %.1f s), and one step on one thread; velocity copied from the GPU flow "
                         "solve" % (cn, p.nx, p.ny, p.nz, threads, cdt)}
    sweep_out = serve_out = None
    if not args.no_sweep:
        sweep_out = sweep_measurement(args)
        serve_out = serving_measurement(args)
    out = {
        "metric": "coupled MLUPS + %HBM roofline; sweep cases/hr and p50/p99 vs plain FP64",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == P.FP64 else "f32",
        "data": "synthetic: seeded sphere pack and biofilm seeds",
        "config": {"workload": w.name, "lattice": [p.nx, p.ny, p.nz], "open_voxels": n_open, "open_interior": n_int,
                   "species": p.n_species, "device": dev,
                   "flow_resolves": int(c1[4] - c0[4]), "ca_sweeps": int(c1[6] - c0[6]),
                   "mask_changes": int(c1[5] - c0[5]), "untimed_steps_to_spill": to_spill},
        "flow": {"resolve_iterations": int(rs_it), "resolve_converged": bool(rs_cv), "resolve_s": round(rs_s, 3),
                 "resolve_interval_steps": int(p.ns_update_interval),
                 "resolve_ms_per_step_amortised": round(resolve_ms_per_step, 4),
                 "mlups_with_a_resolve_every_interval":
                     round(V * args.steps / (t + resolve_ms_per_step * 1e-3 * args.steps) / 1e6, 2),
                 "iteration_us": round(fl_us, 2), "bytes_per_iteration": int(ns_bytes),
                 "iteration_gbs": round(ns_bytes / (fl_us * 1e-6) / 1e9, 1),
                 "iteration_frac": round(ns_bytes / (fl_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                 "note": "k_ns_step + k_ns_tracer per iteration, wall clock over 500 iterations"},
        "units": {"mflups_open": round(n_open * args.steps / t / 1e6, 2),
                  "lattice_mlups": round(V * args.steps * sum(1 for s in range(p.n_species) if p.mobile[s]) / t / 1e6, 2)},
        "roofline": dict(dominant, copy_kernel_gbs=round(copy_gbs, 1)),
        "roofline_transport": transport,
        "roofline_flow": flow,
        "cpu_baseline": cpu,
        "sweep": sweep_out,
        "serving": serve_out,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
