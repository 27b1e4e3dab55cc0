"""Benchmark of the FP64 time step on one MI355X.

Workload: the 3-D periodic turbulence box at 256^3 (Re 1000, Ma 0.3, SL4
derivatives, 10th-order filter every step, turbulenceGenerator forcing),
fixed-dt RK4.  A step is one iteration of the reference time loop: RK4 with
boundary conditions after every stage, the filter in x, y and z, the step
scalars and the state update.

Prints one JSON line.  ``value`` is Mpoint-steps/s with the state resident in
HBM.  ``roofline`` treats the step as the unit of work: the algorithmic bytes
of the minimum-pass schedule (8 B x 377 array touches per point-step in 3-D
without SFD, 8 B x 196 plus the SFD arrays in 2-D) over the duration of one
step, measured with HIP events around replays of the step graph on the
library's stream; ``traffic`` is the HBM bytes one step moves, from rocprofv3
FETCH_SIZE / WRITE_SIZE runs of the same workload (profiles/).  ``kernels``
breaks the step down by kernel class (isolated launches, HIP events).
``other_configs`` gives the 2-D lid-driven cavity (129^2, 1000 steps) and the
2-D boundary layer (1024 x 256) the same way.  ``cpu_baseline`` times the CPU
oracle (a scalar C restatement of the reference path) on a bounded sample of
the same physics.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 Mpoint-steps/s + %HBM roofline; batched-LST & adjoint cost vs plain"
HBM_PEAK_GBPS = 8000.0
MODEL_TOUCHES_3D = 377   # minimum-pass pencil schedule, SFD off
STEP_CLASS = 11          # gatt_time_kernel: one step of the step graph

# measured HBM bytes per step of the 256^3 box (rocprofv3 FETCH_SIZE / WRITE_SIZE runs)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "turbulence_box_256_hbm_bytes.json")

KERNELS = {0: "deriv_x", 1: "deriv_y", 2: "deriv_z", 3: "filter_x", 4: "filter_y", 5: "filter_z",
           6: "visc", 7: "assemble"}
# launches of each class in one step: derivatives 2 per stage, filters 1 per step
LAUNCHES_PER_STEP = {0: 8, 1: 8, 2: 8, 3: 1, 4: 1, 5: 1, 6: 4, 7: 4}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n=64, steps=18):
    from gatt_dns_amd.prep import cases
    from oracle.oracle import Oracle
    case = cases.turbulence_box(n)
    o = Oracle(case)
    o.run(case.state0, 0.0, case.dt, 1)
    t0 = time.perf_counter()
    o.run(case.state0, 0.0, case.dt, steps)
    el = time.perf_counter() - t0
    return {"value": case.npoints * steps / el / 1e6, "unit": "Mpoint-steps/s", "cores": 1, "kind": "port",
            "sample": "CPU oracle (scalar C restatement, gcc -O2, 1 thread) on the same turbulence box at "
                      "%d^3, %d steps, %.1f s" % (n, steps, el)}


def step_roofline(s, reps=5):
    """HIP-event duration of one step of the step graph and the minimum-pass bytes of a step."""
    ms, by = s.time_kernel(STEP_CLASS, reps)
    gbps = by / (ms * 1e-3) / 1e9
    return ms, by, gbps


def plain_config(runtime, label, case, steps, workload):
    """Mpoint-steps/s of one plain run of `steps` steps from the case's initial flow, and the step
    roofline of the same case."""
    s = runtime.Solver(case)
    s.set_state(case.state0)
    s.apply_bound_cond(0.0)
    s.advance(2, 0.0, case.dt)
    s.sync()
    t0 = time.perf_counter()
    mc, dv = s.advance(steps, 2 * case.dt, case.dt)
    s.sync()
    el = time.perf_counter() - t0
    ms, by, gbps = step_roofline(s)
    s.close()
    out = {"workload": workload, "grid": [case.nx, case.ny] + ([case.nz] if case.is3d else []),
           "steps": steps, "value": round(case.npoints * steps / el / 1e6, 3), "unit": "Mpoint-steps/s",
           "ms_per_step": round(el / steps * 1e3, 4), "diverged": bool(dv),
           "step_ms_hip_events": round(ms, 4),
           "model_bytes_per_point_step": round(by / case.npoints, 1),
           "roofline": {"bound": "hbm", "achieved": round(gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                        "frac": round(gbps / HBM_PEAK_GBPS, 4)}}
    log("  %s: %s" % (label, out))
    return out


def batched_throughput(runtime, case, members, steps, shapes=None):
    """Mpoint-steps/s of `members` members advanced together vs one member advanced alone.

    The members start from the case state (gatt_perturb with no vectors), as
    the Krylov members of the stability analysis do; a step of the batch is
    one step of every member.
    """
    res = {}
    for label, nb in (("batched", members), ("single", 1)):
        s = runtime.Solver(case, batch=nb)
        if shapes:
            for (d, k), (v, p, w, _) in shapes.items():
                s.set_line_shape(d, k, v, p, w)
        s.set_base(case.state0)
        s.perturb(0, nb, None, 0.0)
        s.apply_bound_cond(0.0)
        s.advance(2, 0.0, case.dt)
        s.sync()
        t0 = time.perf_counter()
        s.advance(steps, 2 * case.dt, case.dt)
        s.sync()
        el = time.perf_counter() - t0
        res[label] = case.npoints * nb * steps / el / 1e6
        s.close()
        del s
    return {"members": members, "steps": steps, "batched_mpts": round(res["batched"], 2),
            "single_mpts": round(res["single"], 2), "speedup_vs_sequential": round(res["batched"] / res["single"], 3)}


def ragged_sweep(runtime, steps=20):
    """16 lid-driven cavities 256x128 .. 768x256, Re 500..2000: grouped vs one after the other."""
    from gatt_dns_amd.prep import cases
    sweep = cases.cavity_sweep()
    pts = sum(c.npoints for c in sweep)

    def fresh():
        out = []
        for c in sweep:
            s = runtime.Solver(c)
            s.set_state(c.state0)
            s.apply_bound_cond(0.0)
            out.append(s)
        return out

    res = {}
    solvers = fresh()
    for s, c in zip(solvers, sweep):
        s.advance(2, 0.0, c.dt)
    for s in solvers:
        s.sync()
    t0 = time.perf_counter()
    for s, c in zip(solvers, sweep):
        s.advance(steps, 2 * c.dt, c.dt)
        s.sync()
    res["sequential"] = pts * steps / (time.perf_counter() - t0) / 1e6
    for mode in (0, 1, 2):
        g = runtime.Group(solvers, mode=mode)
        g.advance(2, 0.0, [c.dt for c in sweep])
        t0 = time.perf_counter()
        g.advance(steps, 0.0, [c.dt for c in sweep])
        for s in solvers:
            s.sync()
        res["group_mode%d" % mode] = pts * steps / (time.perf_counter() - t0) / 1e6
        g.close()
    for s in solvers:
        s.close()
    return {"members": len(sweep), "points": pts, "steps": steps,
            "sequential_mpts": round(res["sequential"], 2),
            "group_graph_mpts": round(res["group_mode0"], 2),
            "group_streams_mpts": round(res["group_mode1"], 2),
            "group_ragged_launch_mpts": round(res["group_mode2"], 2),
            "speedup_vs_sequential": round(max(res["group_mode0"], res["group_mode1"], res["group_mode2"])
                                           / res["sequential"], 3)}


def adjoint_cost(runtime, nx=512, ny=192, nz=64, steps=6, only_adjoint=False, base_steps=10000, long_steps=2000,
                 every=100):
    """Tangent and adjoint step cost against the plain step on the adjoint configuration: the
    spanwise-periodic boundary layer 512x192x64 (the 2-D boundary layer extruded in z; walls, buffers,
    SFD with a frozen mean in the buffers) about a steady base (an SFD run of the 2-D boundary layer,
    replicated in z); then the wall time of `long_steps` forward steps and of the gradient of
    E = 1/2 ||q'(T)||^2 over `long_steps` linearised steps with checkpoints every `every` steps."""
    import torch
    from gatt_dns_amd.adjoint import Propagator, optimal_energy_gradient, state_tensor
    from gatt_dns_amd.baseflow import extrude, sfd_base
    from gatt_dns_amd.prep import cases
    # (the profiled form times steps only: a short base run)
    _, q2, base_info = sfd_base(nx, ny, steps=500 if only_adjoint else base_steps)
    case = cases.boundary_layer(nx=nx, ny=ny, nz=nz, noise=0.0, sfd_delta=-1.0)
    prop = Propagator(case)
    q0 = state_tensor(case, extrude(q2, nz))
    v = torch.randn(prop.shape, dtype=torch.float64, device="cuda")
    res = {}
    prop.load(q0)
    prop.solver.apply_bound_cond(0.0)
    q0 = prop.save()

    def measure(tag):
        prop.load(q0)
        prop.steps(2)
        prop.solver.sync()
        t0 = time.perf_counter()
        prop.steps(steps)
        prop.solver.sync()
        res["forward" + tag] = (time.perf_counter() - t0) / steps
        prop.load(q0)
        prop.load_perturbation(v)
        prop.solver.tlm_step(0.0, prop.dt)
        prop.solver.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            prop.solver.tlm_step(0.0, prop.dt)
        prop.solver.sync()
        res["tangent" + tag] = (time.perf_counter() - t0) / steps
        prop.load(q0)
        prop.load_perturbation(v)
        prop.solver.adj_step(0.0, prop.dt)
        t0 = time.perf_counter()
        for _ in range(steps):
            prop.solver.adj_step(0.0, prop.dt)
        prop.solver.sync()
        res["adjoint" + tag] = (time.perf_counter() - t0) / steps

    if only_adjoint:
        # one warm-up adjoint step, then `steps` timed ones (the profiled form)
        prop.load(q0)
        prop.load_perturbation(v)
        prop.solver.adj_step(0.0, prop.dt)
        prop.solver.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            prop.solver.adj_step(0.0, prop.dt)
        prop.solver.sync()
        ms = (time.perf_counter() - t0) / steps * 1e3
        prop.close()
        return {"workload": "spanwise-periodic boundary layer %dx%dx%d adjoint steps" % (nx, ny, nz),
                "steps": steps, "adjoint_ms_per_step": round(ms, 3)}
    measure("")
    # the long run: forward steps, then the checkpointed gradient of the perturbation energy
    prop.load(q0)
    prop.solver.sync()
    t0 = time.perf_counter()
    prop.steps(long_steps)
    prop.solver.sync()
    long_fwd = time.perf_counter() - t0
    dq0 = v / torch.linalg.norm(v)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    energy, grad = optimal_energy_gradient(prop, q0, dq0, long_steps, checkpoint_every=every)
    torch.cuda.synchronize()
    long_grad = time.perf_counter() - t0
    del grad
    torch.cuda.empty_cache()
    # the same with the line-solve shapes picked by the autotuner for this case
    from gatt_dns_amd import autotune
    autotune.tune(prop.solver, reps=3)
    measure("_tuned")
    prop.close()
    return {"workload": "spanwise-periodic boundary layer %dx%dx%d, Re 1000, Ma 0.5, walls, buffers, "
                        "SFD (frozen mean) in the buffers, about a steady base (2-D SFD run of %d steps, "
                        "chi 0.1, Delta 5, last max change %.2e, replicated in z)"
                        % (nx, ny, nz, base_info["steps"], max(base_info["max_change"])),
            "long_run": {"steps": long_steps, "checkpoint_every": every, "forward_s": round(long_fwd, 3),
                         "gradient_s": round(long_grad, 3), "gradient_over_forward": round(long_grad / long_fwd, 3),
                         "note": "gradient = tangent run keeping checkpoints + per segment recompute of the base "
                                 "states and the adjoint steps in reverse", "energy": energy},
            "forward_ms_per_step": round(res["forward"] * 1e3, 3),
            "tangent_ms_per_step": round(res["tangent"] * 1e3, 3),
            "adjoint_ms_per_step": round(res["adjoint"] * 1e3, 3),
            "adjoint_over_forward": round(res["adjoint"] / res["forward"], 3),
            "tuned_shapes": {"forward_ms_per_step": round(res["forward_tuned"] * 1e3, 3),
                             "tangent_ms_per_step": round(res["tangent_tuned"] * 1e3, 3),
                             "adjoint_ms_per_step": round(res["adjoint_tuned"] * 1e3, 3),
                             "adjoint_over_forward": round(res["adjoint_tuned"] / res["forward_tuned"], 3)},
            "note": "adjoint step includes recomputing its base step's stage states and boundary applications"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-lst", action="store_true", help="skip the batched-member (stability analysis) lines")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the tangent/adjoint cost line")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only time --steps adjoint steps of the boundary layer (prints their JSON; for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (reference for HBM byte counters)")
    ap.add_argument("--no-other", action="store_true", help="skip the cavity and boundary-layer lines")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run this many steps of the box with the default shapes (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the time-advance path runs on a single MI355X here; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from gatt_dns_amd import runtime
    from gatt_dns_amd.prep import cases

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.adjoint_only:
        print(json.dumps(adjoint_cost(runtime, steps=args.steps, only_adjoint=True)), flush=True)
        return
    if args.profile_steps:
        # the profiled form: nothing but the steps (one graph of two steps replayed)
        case = cases.turbulence_box(args.n)
        s = runtime.Solver(case)
        s.set_state(case.state0)
        s.sync()
        t0 = time.perf_counter()
        s.advance(args.profile_steps, 0.0, case.dt)
        s.sync()
        el = time.perf_counter() - t0
        s.close()
        if args.calibrate_copy:
            src = torch.rand(1 << 25, dtype=torch.float64, device="cuda")
            dst = torch.empty_like(src)
            for _ in range(3):
                dst.copy_(src)
            torch.cuda.synchronize()
        print(json.dumps({"profile_steps": args.profile_steps, "grid": [args.n] * 3,
                          "ms_per_step": round(el / args.profile_steps * 1e3, 4)}), flush=True)
        return
    log("preparing turbulence box %d^3" % args.n)
    case = cases.turbulence_box(args.n)
    s = runtime.Solver(case)
    tuned = None
    if not args.no_tune:
        from gatt_dns_amd import autotune
        log("autotuning line-solve shapes")
        tuned = autotune.tune(s, reps=3, log=log)
        for (d, k), (v, p, w, ms) in sorted(tuned.items()):
            log("  dir %d kind %d -> variant %d param %d wpb %d, %.4f ms" % (d, k, v, p, w, ms))
    s.set_state(case.state0)
    s.apply_bound_cond(0.0)
    dt = case.dt
    log("device: %s; dt = %.6g; warmup %d" % (runtime.device_info(), dt, args.warmup))
    if args.warmup:
        s.advance(args.warmup, 0.0, dt)
    s.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mc, dv = s.advance(args.steps, args.warmup * dt, dt)
    s.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_step = el / args.steps * 1e3
    value = case.npoints * args.steps / el / 1e6
    log("timed %d steps: %.3f ms/step, %.1f Mpoint-steps/s, max changes %s, diverged %s"
        % (args.steps, ms_step, value, mc, dv))

    # per-kernel-class timing with HIP events on the library stream
    per_class = {}
    for k, name in KERNELS.items():
        if k in (3, 4, 5):
            continue
        ms, by = s.time_kernel(k, 10)
        per_class[name] = {"ms": ms, "bytes": by, "share_ms_per_step": ms * LAUNCHES_PER_STEP[k]}
    for k in (3, 4, 5):   # filters overwrite the state: time them last
        ms, by = s.time_kernel(k, 10)
        per_class[KERNELS[k]] = {"ms": ms, "bytes": by, "share_ms_per_step": ms * LAUNCHES_PER_STEP[k]}
    for name, v in per_class.items():
        log("  %-9s %.4f ms/launch  %.0f GB/s algorithmic  %.3f ms/step"
            % (name, v["ms"], v["bytes"] / (v["ms"] * 1e-3) / 1e9, v["share_ms_per_step"]))
    shapes_out = {"%s_%s" % ("xyz"[d], "deriv" if k == 0 else "filter"): list(s.get_line_shape(d, k))
                  for d in range(3) for k in (0, 1)}
    # the step as the unit: minimum-pass bytes over the HIP-event duration of one graph-replayed step
    s.set_state(case.state0)
    s.advance(2, 0.0, dt)
    step_ms, step_bytes, step_gbps = step_roofline(s)
    log("step graph: %.4f ms/step (HIP events), %.0f GB/s of minimum-pass bytes" % (step_ms, step_gbps))
    # measured HBM bytes of one step (rocprofv3 FETCH_SIZE / WRITE_SIZE runs of the same workload,
    # calibrated against a device copy; committed under profiles/)
    traffic = None
    if args.n == 256 and os.path.exists(TRAFFIC_FILE):
        traffic = json.load(open(TRAFFIC_FILE)).get("bytes_per_step")

    other = None
    if not args.no_other:
        s.close()
        del s
        torch.cuda.empty_cache()
        other = {}
        log("lid-driven cavity 129^2")
        other["lid_driven_129"] = plain_config(
            runtime, "lid_driven_129", cases.lid_driven(129), 1000,
            "2-D lid-driven cavity 129x129, Re 1000, Ma 0.5, fluid at rest, lid u = 1, fixed-dt RK4 (CFL 0.5), "
            "filter every step")
        log("boundary layer 1024x256")
        other["boundary_layer_1024x256"] = plain_config(
            runtime, "boundary_layer_1024x256", cases.boundary_layer(), 200,
            "2-D isothermal-wall boundary layer 1024x256 (x in [-1, 30], y in [0, 8], wall ratio 1.02, "
            "40-point buffers), Re 1000, Ma 0.5, compressible similarity initial flow + 1e-5 noise, "
            "SFD chi 0.02, Delta 10")
        s = None

    if args.calibrate_copy:
        src = torch.rand(1 << 25, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 3 x %d bytes read + written" % (src.numel() * 8))

    lst = None
    if not args.no_lst:
        # stability analysis: a block of 32 Krylov members advanced together versus
        # 32 sequential plain runs (calcInstability.m launches one DNS per vector)
        lst = {}
        if s is not None:
            s.close()
            s = None
        torch.cuda.empty_cache()
        for label, mk, nsteps, shapes in (
                ("turbulence_box_%d" % args.n, lambda: case, 2, tuned),
                ("turbulence_box_64", lambda: cases.turbulence_box(64), 10, None),
                ("lid_driven_129", lambda: cases.lid_driven(129), 50, None)):
            c2 = mk()
            log("batched members: %s" % label)
            lst[label] = batched_throughput(runtime, c2, 32, nsteps, shapes)
            log("  %s" % lst[label])
        log("ragged sweep: 16 lid-driven cavities")
        lst["ragged_cavity_sweep"] = ragged_sweep(runtime)
        log("  %s" % lst["ragged_cavity_sweep"])

    adj = None
    if not args.no_adjoint:
        log("tangent / adjoint step cost")
        adj = adjoint_cost(runtime)
        log("  %s" % adj)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle sample)")
        cpu = cpu_baseline()

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpoint-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded random-phase solenoidal velocity, E(k)~k^4 exp(-2(k/4)^2), u_rms 0.5)",
        "config": {"workload": "3-D periodic turbulence box %d^3, Re 1000, Ma 0.3, SL4 + 10th-order filter "
                               "(alpha 0.49) every step, turbulenceGenerator forcing, fixed-dt RK4" % args.n,
                   "grid": [args.n] * 3, "dt": dt, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "rk4_step_graph", "achieved": round(step_gbps, 1),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(step_gbps / HBM_PEAK_GBPS, 4),
                     "traffic": traffic,
                     "traffic_source": "profiles/turbulence_box_256_hbm_bytes.json" if traffic else None,
                     "model_bytes_per_point_step": 8 * MODEL_TOUCHES_3D, "bytes_per_step": step_bytes,
                     "ms_per_step_hip_events": round(step_ms, 5)},
        "kernels": {k: {"ms_per_launch": round(v["ms"], 5), "ms_per_step": round(v["share_ms_per_step"], 4),
                        "algorithmic_gbps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)}
                    for k, v in per_class.items()},
        "line_shapes": shapes_out,
        "other_configs": other,
        "batched_lst": lst,
        "adjoint": adj,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
