#!/usr/bin/env python3
"""The FP32 chosen form against the FP64 plain form on one workload, from
the same initial state:

- instantaneous fields after `--inst-steps` steps: the largest difference of
  each prognostic field relative to the FP64 field's largest magnitude;
- time means over `--hours` of model time on the plane 20 m above the vent
  (1 Hz samples, the statistics path of hc_stats_*): relative L2 differences
  of the mean w (upward) and T-anomaly profiles across the plume (the plane's
  means along y and along x) and of the TKE and w'T' profiles;
- the same differences between two FP64 runs whose 1e-4 K initial noise
  differs only in its seed: the spread the time means have anyway, since LES
  trajectories decorrelate.

usage: python compare_forms.py [--workload les384] [--hours 1.0] [--inst-steps 50]
Prints one JSON line; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import hotcross as H  # noqa: E402

FIELDS = ("u", "v", "w", "p", "t", "s", "c", "rhop")


def context(name, precision, seed):
    p = H.workload_params(name)
    ctx = H.Context(p, precision)
    z, t, s = H.linear_stratification(p)
    ctx.init_case(z, t, s, t_noise=1.0e-4, seed=seed)
    return p, ctx


def instantaneous(name, steps):
    out = {}
    for prec in (H.FP64, H.FP32):
        _, ctx = context(name, prec, 2024)
        ctx.step(steps)
        out[prec] = {f: ctx.download(f) for f in FIELDS}
        ctx.close()
    errs = {}
    for f in FIELDS:
        a, b = out[H.FP32][f], out[H.FP64][f]
        m = np.abs(b) < 1e30
        errs[f] = float(np.max(np.abs(a[m] - b[m])) / max(np.max(np.abs(b[m])), 1e-300))
    return errs


def plane_means(name, precision, seed, hours):
    p, ctx = context(name, precision, seed)
    try:
        kp = p.nz - int(round(20.0 / p.dz))
        t0 = ctx.download("t")[1:p.nx + 1, 1:p.ny + 1, kp].T.copy()
        every = max(1, int(round(1.0 / p.dt)))
        ctx.stats_reset(kp, kp, every)
        n = int(round(hours * 3600.0 / p.dt))
        chunk = max(every, 300 // every * every)
        done, tic = 0, time.time()
        while done < n:
            c = min(chunk, n - done)
            ctx.step(c)
            done += c
            ctx.synchronize()
            sys.stderr.write("precision %d seed %d: %d/%d steps, %.0f s\n" % (precision, seed, done, n,
                                                                          time.time() - tic))
            sys.stderr.flush()
        st = ctx.stats()
        return {"w_up": st["w_up"][0], "t_anom": st["t"][0] - t0, "tke": st["tke"][0], "wt_flux": st["wt_flux"][0]}
    finally:
        ctx.close()


def profile_errors(a, b):
    rel = lambda x, y: float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
    out = {}
    for k in a:
        out[k + "_across_x"] = rel(a[k].mean(axis=0), b[k].mean(axis=0))
        out[k + "_across_y"] = rel(a[k].mean(axis=1), b[k].mean(axis=1))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="les384")
    ap.add_argument("--hours", type=float, default=1.0)
    ap.add_argument("--inst-steps", type=int, default=50)
    args = ap.parse_args()
    res = {"workload": args.workload, "inst_steps": args.inst_steps, "hours": args.hours}
    res["instantaneous_rel_max_err"] = instantaneous(args.workload, args.inst_steps)
    sys.stderr.write("instantaneous: %s\n" % res["instantaneous_rel_max_err"])
    a64 = plane_means(args.workload, H.FP64, 2024, args.hours)
    a32 = plane_means(args.workload, H.FP32, 2024, args.hours)
    b64 = plane_means(args.workload, H.FP64, 2025, args.hours)
    res["time_mean_fp32_vs_fp64"] = profile_errors(a32, a64)
    res["time_mean_fp64_seed_spread"] = profile_errors(b64, a64)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
