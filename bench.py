#!/usr/bin/env python3
"""Headline measurement: Mcell-steps/s of the HotCross LES timestep on one
MI355X (FP64 plain form, 256^3 workload), with the dominant kernel's HBM
roofline and the reference's OpenMP CPU path timed beside it.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload les256]
                [--precision fp64|fp32] [--no-cpu] [--no-fp32] [--no-adjoint]
                [--only-adjoint] [--adjoint-workload les128] [--adjoint-steps 400]
                [--adjoint-seg 20] [--adjoint-precision fp64] [--adjoint-budget-gb G]
                [--c5-steps 1800] [--c5-budget-gb 200]

A step is one pass of the reference's time loop body
(hotcross80-revised2014.f:264-350: body force, forces, pressure solve,
momentum, Smagorinsky viscosity, source, three MPDATA tracers, density,
Asselin/leapfrog refresh, CFL every 10 steps) over the whole grid, with the
state resident in HBM.  Mcell-steps/s = nx*ny*nz*K / seconds.

The same line carries the FP32 form's throughput on the same workload and its
error against the FP64 form after 50 steps from the same state ("fp32"), and
the adjoint's cost: "adjoint" is the 128^3 FP64 adjoint configuration (400
steps, a checkpoint every 20, 25 probes at 1 Hz); "adjoint_256" the 256^3
FP32 forward + adjoint configuration (1800 steps, the checkpointing
plan_window chooses under 200 GB: a fixed interval where one fits, binomial
checkpointing otherwise).  Each window is timed beside the same
number of forward steps of the step body the window runs (T and S, no passive
tracer), and reports the reverse sweep per forward step, (window - forward) /
forward, and the whole window per forward step.
"""
import argparse
import ctypes
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "plumemodel-hotcross_amd"))

HBM_PEAK = 8.0e12          # MI355X HBM3E, B/s
# algorithmic HBM bytes per cell-step: 130.3 field moves (compulsory reads and
# writes of every reference subroutine, counted once each) x 8 B in FP64
FIELD_MOVES = 130.3
# MPDATA corrective pass, one launch of k_mpdata_pass3: pass 2 reads c and the
# three Courant numbers and writes the three antidiffusive velocities and the
# result (8 field moves per cell); pass 3 reads c, the three antidiffusive
# velocities and the field's entry value and writes the result (6 moves).
# The mean over a tracer's two launches is 7 moves per cell.
MPDATA_PASS_MOVES = 7.0


def profiled_traffic(workload, precision):
    """HBM bytes per launch of k_mpdata_pass3 from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload (None if there is none)."""
    path = os.path.join(ROOT, "profiles", "k_mpdata_pass3_%s_%s_hbm.md" % (workload, precision))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch ="):
            return float(line.split("=", 1)[1])
    return None


def init_context(H, name, precision):
    p = H.workload_params(name)
    ctx = H.Context(p, precision)
    z, t, s = H.linear_stratification(p)
    ctx.init_case(z, t, s, t_noise=1.0e-4, seed=2024)
    return p, ctx


def gpu_leg(args, precision=None, workload=None):
    import hotcross as H
    precision = precision or args.precision
    workload = workload or args.workload
    prec = H.FP64 if precision == "fp64" else H.FP32
    p, ctx = init_context(H, workload, prec)
    ctx.step(args.warmup)
    ctx.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0
    cells = p.nx * p.ny * p.nz
    # phase and kernel timing with HIP events on the context stream (separate
    # run so the events do not perturb the timed region)
    ctx.phase_timing(True)
    nprof = min(args.steps, 10)
    ctx.step(nprof)
    ctx.synchronize()
    phases = ctx.phase_times()
    ctx.phase_timing(False)
    ctx.close()
    return p, sec, cells, phases, nprof


def adjoint_leg(workload="les128", nsteps=400, seg=20, precision="fp64", budget=None):
    """The adjoint window of an adjoint configuration: nsteps steps of the
    timestep (FP64 unless precision is fp32), 25 probes on the plane 20 m
    above the vent sampled at 1 Hz, a checkpoint every seg steps (or the
    interval plan_window chooses under `budget` bytes, or binomial
    checkpointing with as many snapshots as fit where no interval does),
    then the reverse sweep (segments recomputed into an HBM tape, adjoint of
    every step).
    Timed beside the same number of forward steps (hc_step) of the body the
    window runs: T and S without the passive tracer, which no loss reads."""
    import hotcross as H
    import adjoint as AD
    p = H.workload_params(workload)
    p.toggle_c = 0  # the window's forward steps advance T and S only
    ctx = H.Context(p, H.FP64 if precision == "fp64" else H.FP32)
    z, t, s = H.linear_stratification(p)
    ctx.init_case(z, t, s, t_noise=1.0e-4, seed=2024)
    ctx.step(2)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(nsteps)
    ctx.synchronize()
    fwd = time.perf_counter() - t0
    need, kind = None, "fixed"
    if budget:
        kind, seg, need = AD.plan_window(ctx, nsteps, budget)
    probes = AD.plane_probes(p)
    every = AD.samples_every(p)
    ctx.set_probes(probes, every, 1.0e4, 1.0)
    obs = np.zeros((nsteps // every, len(probes), 2))
    t0 = time.perf_counter()
    if kind == "fixed":
        loss, _, gq, gs = ctx.adj_window(nsteps, seg, obs)
    else:
        loss, _, gq, gs = ctx.adj_window_binomial(nsteps, seg, obs)
    ctx.synchronize()
    win = time.perf_counter() - t0
    ctx.close()
    how = ("checkpoint every %d" % seg) if kind == "fixed" else ("binomial checkpointing, %d snapshot slots, %d "
                                                                "forward steps" % (seg, H.binomial_schedule(nsteps, seg)[1]))
    out = {"workload": "%s %s, %d steps, %s, %d probes at 1 Hz" % (workload, precision, nsteps, how, len(probes)),
           "checkpointing": kind,
           "forward_ms_per_step": round(fwd / nsteps * 1e3, 4),
           "window_ms_per_step": round(win / nsteps * 1e3, 4),
           "ratio": round((win - fwd) / fwd, 3),
           "window_over_forward": round(win / fwd, 3),
           "note": "forward = hc_step on T and S (the window's step body, CFL every 10 steps); ratio = (window - "
                   "forward) / forward: the reverse sweep (segment recompute + adjoint of every step) per forward "
                   "step; window_over_forward includes the window's own forward sweep",
           "finite": bool(np.isfinite(loss) and np.all(np.isfinite(gq)) and np.all(np.isfinite(gs)))}
    if need is not None:
        out["hbm_bytes"] = need
        out["budget_bytes"] = budget
    return out


def step_only_leg(workload, nsteps, precision):
    """The adjoint of the timestep alone: a window of nsteps steps whose one
    segment the forward sweep tapes as it runs, so that the reverse sweep
    recomputes nothing and runs only the adjoint of each step.  (window -
    forward) / forward is then the adjoint of one step (plus the taping in
    the forward sweep) per forward step."""
    r = adjoint_leg(workload, nsteps, nsteps, precision)
    return {"workload": "%s %s, %d steps, one segment taped by the forward sweep (no recompute)"
                        % (workload, precision, nsteps),
            "forward_ms_per_step": r["forward_ms_per_step"], "window_ms_per_step": r["window_ms_per_step"],
            "adjoint_step_over_forward": r["ratio"], "finite": r["finite"]}


def fp32_error_leg(workload, nsteps=50):
    """The FP32 form against the FP64 form after nsteps steps of the native
    loop from the same initial state: per field, the largest difference
    relative to the field's largest FP64 magnitude."""
    import hotcross as H
    p = H.workload_params(workload)
    z, t, s = H.linear_stratification(p)
    ctxs = []
    for prec in (H.FP64, H.FP32):
        c = H.Context(p, prec)
        c.init_case(z, t, s, t_noise=1.0e-4, seed=2024)
        c.step(nsteps)
        ctxs.append(c)
    err = {}
    for f in ("u", "v", "w", "p", "t", "s", "c", "ss", "ss_v", "rhop"):
        a, b = ctxs[0].download(f), ctxs[1].download(f)
        m = np.abs(a) < 1e30
        err[f] = float("%.3g" % (np.abs(a[m] - b[m]).max() / max(np.abs(a[m]).max(), 1e-300)))
    for c in ctxs:
        c.close()
    return err


def host_cpus():
    """(model name, usable logical CPUs, usable physical cores) of this host:
    the affinity mask this process may run on, and the distinct
    (package, core) pairs among those CPUs."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = "/sys/devices/system/cpu/cpu%d/topology/" % c
        try:
            cores.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(cpus), len(cores)


def cpu_leg(nx, ny, nz, steps, threads, phases=False):
    """The reference compiled in place (oracle/_ref) on this host's cores
    (with phases, the worker reports each step on stderr, which passes
    through)."""
    lib = os.path.join(ROOT, "oracle", "_ref", f"g{nx}x{ny}x{nz}", "libhcref.so")
    if not os.path.exists(lib):
        return None
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    env.setdefault("OMP_PROC_BIND", "close")
    env.setdefault("OMP_PLACES", "cores")
    env.setdefault("OMP_STACKSIZE", "512M")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", lib, str(nx), str(ny), str(nz), str(steps),
           str(threads), "1" if phases else "0"]

    def big_stack():
        # the reference keeps grid-sized automatic arrays on the main
        # thread's stack (OMP_STACKSIZE covers only the worker threads)
        import resource
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run(cmd, cwd=d, env=env, stdout=subprocess.PIPE, stderr=None if phases else subprocess.PIPE,
                           text=True, timeout=1800, preexec_fn=big_stack)
    if r.returncode != 0:
        if r.stderr:
            sys.stderr.write(r.stderr[-2000:])
        return None
    # the reference's own Fortran output shares stdout; the worker's JSON
    # line is the one that starts with '{'
    lines = [x for x in r.stdout.splitlines() if x.startswith("{")]
    return json.loads(lines[-1]) if lines else None


# the reference's time loop body (hotcross80-revised2014.f:264-350) as the
# harness calls it, one subroutine at a time, for per-phase timing
REF_PHASES = ("body_force", "forces", "pressure", "momentum_u", "momentum_v", "momentum_w", "viscosity3d",
              "source_terms", "mixing", "work_velocity", "tracer_t", "tracer_s", "density", "tracer_c", "refresh")


def cpu_worker(lib, nx, ny, nz, steps, threads, phases=False):
    """Runs inside a fresh process: set up a case in the reference's module
    state and time `steps` steps of its time loop body on `threads` OpenMP
    threads (with phases, each subroutine call timed on its own; the step is
    then the same calls the harness's step makes, in the same order)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import hc_testlib as T
    d = 1.0 if nx <= 64 else 0.5
    prm = T.reference_params(nx, ny, nz, dx=d, dy=d, dz=d, dt=0.25 if nx <= 64 else 0.2, bit7=0)
    ref = T.RefLib(lib, prm, threads=threads)
    z, t, s = T.linear_profile_table(zmin=prm.zstart - 10, zmax=prm.zend + 10, dzt=1.0)
    T.write_profile_file("profile.txt", z, t, s)
    ref.init_profiles(os.path.abspath("profile.txt"))
    ref.call("density"); ref.call("initialize"); ref.call("sponge_maker")
    ref.call("step", 0)  # first step (allocates the forcing series)
    ph = dict.fromkeys(REF_PHASES, 0.0)
    t0 = time.perf_counter()
    for m in range(1, steps + 1):
        if not phases:
            ref.call("step", m)
            continue
        ref.call("begin_step", m)
        for name in REF_PHASES:
            a = time.perf_counter()
            if name.startswith("tracer_"):
                ref.call("tracer", "tsc".index(name[-1]) + 1)
            else:
                ref.call(name)
            ph[name] += time.perf_counter() - a
        print("  step %d of %d, %.1f s" % (m, steps, time.perf_counter() - t0), file=sys.stderr, flush=True)
    sec = time.perf_counter() - t0
    out = {"sec": sec, "steps": steps, "threads": threads}
    if phases:
        out["phases_s_per_step"] = {k: round(v / steps, 5) for k, v in ph.items()}
    print(json.dumps(out), flush=True)


def cpu_full(ncore):
    """The whole CPU baseline plan: the reference at 64^3 (dx = 1 m, dt =
    0.25 s, 200 steps) and at 256^3 (dx = 0.5 m, dt = 0.2 s, 10 steps), on
    all usable physical cores and on 8 threads (the reference's own setting,
    hotcross80-revised2014.f:88-90), with per-subroutine times."""
    rows = []
    for (nx, steps) in ((64, 200), (256, 10)):
        for threads in sorted({ncore, 8}):
            print("cpu baseline %d^3, %d steps, %d threads" % (nx, steps, threads), file=sys.stderr, flush=True)
            r = cpu_leg(nx, nx, nx, steps, threads, phases=True)
            if r:
                r["grid"] = [nx, nx, nx]
                r["mcell_steps_per_s"] = round(nx ** 3 * steps / r["sec"] / 1e6, 4)
            rows.append(r)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="les256")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 form's line")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint window timing")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="time only the adjoint window (prints its JSON object), e.g. for a profiler")
    ap.add_argument("--adjoint-workload", default="les128")
    ap.add_argument("--adjoint-steps", type=int, default=400)
    ap.add_argument("--adjoint-seg", type=int, default=20)
    ap.add_argument("--adjoint-precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--adjoint-budget-gb", type=float, default=0.0,
                    help="with --only-adjoint: plan the checkpoints under this HBM budget (GB) instead of "
                         "--adjoint-seg")
    ap.add_argument("--c5-steps", type=int, default=1800, help="steps of the 256^3 FP32 adjoint window")
    ap.add_argument("--c5-budget-gb", type=float, default=200.0,
                    help="HBM budget of the 256^3 FP32 adjoint window's checkpoints (GB)")
    ap.add_argument("--cpu-worker", nargs=7)
    ap.add_argument("--cpu-full", action="store_true",
                    help="only the whole CPU baseline plan (64^3 and 256^3, all physical cores and 8 threads)")
    args = ap.parse_args()
    if args.cpu_worker:
        lib, nx, ny, nz, steps, threads, phases = args.cpu_worker
        cpu_worker(lib, int(nx), int(ny), int(nz), int(steps), int(threads), phases == "1")
        return
    if args.cpu_full:
        model, ncpu, ncore = host_cpus()
        print(json.dumps({"host": {"model": model, "usable_cpus": ncpu, "usable_physical_cores": ncore},
                          "runs": cpu_full(ncore)}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: the timestep has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.only_adjoint:
        budget = args.adjoint_budget_gb * 1e9 if args.adjoint_budget_gb else None
        print(json.dumps(adjoint_leg(args.adjoint_workload, args.adjoint_steps, args.adjoint_seg,
                                     args.adjoint_precision, budget=budget)))
        return
    p, sec, cells, phases, nprof = gpu_leg(args)
    value = cells * args.steps / sec / 1e6
    ms_step = sec * 1e3 / args.steps
    esz = 8 if args.precision == "fp64" else 4
    # roofline of the dominant kernel: the fused MPDATA corrective pass
    # (k_mpdata_pass3), timed per launch with HIP events on the context stream
    roof = None
    k_ms, k_n = phases.get("mpdata_pass", (0.0, 0))
    if k_n:
        per_launch = k_ms / k_n * 1e-3
        alg = MPDATA_PASS_MOVES * esz * cells
        achieved = alg / per_launch / 1e9
        roof = {"bound": "hbm", "kernel": "k_mpdata_pass3 (MPDATA corrective pass, mean of passes 2 and 3)",
                "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved / (HBM_PEAK / 1e9), 4),
                "traffic": profiled_traffic(args.workload, args.precision),
                "avg_ms": round(per_launch * 1e3, 4), "launches": k_n,
                "alg_bytes_per_launch": alg}
    step_bytes = FIELD_MOVES * esz * cells
    fp32 = None
    if args.precision == "fp64" and not args.no_fp32:
        p32, sec32, _, ph32, n32 = gpu_leg(args, "fp32")
        v32 = cells * args.steps / sec32 / 1e6
        fp32 = {"value": round(v32, 3), "unit": "Mcell-steps/s", "ms_per_step": round(sec32 * 1e3 / args.steps, 4),
                "speedup_vs_fp64": round(v32 / value, 3), "dtype": "f32 (FP64 islands)",
                "err_vs_fp64": fp32_error_leg(args.workload),
                "err_note": "per field, max |FP32 - FP64| / max |FP64| after 50 native steps from the same state"}
    adj = adj256 = None
    if not args.no_adjoint:
        adj = adjoint_leg(args.adjoint_workload, args.adjoint_steps, args.adjoint_seg, args.adjoint_precision)
        adj["step_only"] = step_only_leg(args.adjoint_workload, 20, args.adjoint_precision)
        # the headline size's forward + adjoint configuration: 256^3 in the
        # FP32 form, 1800 steps, checkpoints under a 200 GB HBM budget
        adj256 = adjoint_leg("les256", args.c5_steps, None, "fp32", budget=args.c5_budget_gb * 1e9)
        adj256["step_only"] = step_only_leg("les256", 10, "fp32")
    # the CPU baseline last: run right after it, the launch-bound 128^3
    # adjoint leg once measured 1.5x slower
    cpu = None
    if not args.no_cpu:
        model, ncpu, ncore = host_cpus()
        threads = int(os.environ.get("HC_CPU_THREADS", ncore))
        r = cpu_leg(p.nx, p.ny, p.nz, args.cpu_steps, threads)
        if r:
            cpu = {"value": round(cells * r["steps"] / r["sec"] / 1e6, 4), "unit": "Mcell-steps/s",
                   "cores": threads, "kind": "reference",
                   "host": {"model": model, "usable_cpus": ncpu, "usable_physical_cores": ncore},
                   "sample": "%d steps of the reference's time loop body at %dx%dx%d (amdflang -O3 OpenMP, "
                             "reference constants, OMP_NUM_THREADS=%d on the usable physical cores), after one "
                             "untimed step" % (r["steps"], p.nx, p.ny, p.nz, threads)}
    out = {
        "metric": "Mcell-steps/s FP64 vs FP32 at stated err vs FP64; adjoint/forward time ratio",
        "value": round(value, 3), "unit": "Mcell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if esz == 8 else "f32",
        "data": "synthetic (linear stratification N=1e-3 1/s, seeded T noise, 50 MW vent)",
        "config": {"workload": args.workload, "grid": [p.nx, p.ny, p.nz], "dx": p.dx, "dt": p.dt,
                   "precision": args.precision, "parallelism": "single GPU"},
        "step_roofline": {"bytes_per_cell_step": FIELD_MOVES * esz,
                          "achieved_GBps": round(step_bytes / (ms_step * 1e-3) / 1e9, 1),
                          "frac": round(step_bytes / (ms_step * 1e-3) / HBM_PEAK, 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp32": fp32,
        "adjoint": adj,
        "adjoint_256": adj256,
        "phases_ms_per_step": {k: round(v[0] / nprof, 4) for k, v in phases.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
