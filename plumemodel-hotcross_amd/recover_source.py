#!/usr/bin/env python3
"""Recovery of the vent's heat-flux map from probe data: the source-recovery
configuration of the adjoint (256^3, dx = 0.5 m, 1800 steps, 1 Hz w/T probes
on the plane 20 m above the vent, checkpoints sized to an HBM budget).

A seeded "true" map (a per-cell factor on the vent heat flux, mean 1 so the
total stays 50 MW) produces synthetic probe data; L-BFGS through the torch
harness (adjoint.Window) then fits the map from a uniform start.  Each
evaluation is one adjoint window (forward with checkpoints, reverse sweep).

usage: python recover_source.py [--workload les256] [--precision fp32]
           [--steps 1800] [--spinup 300] [--evals 8] [--budget-gb 200]
Prints one line per evaluation (loss, map error) and a JSON summary.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import adjoint as AD  # noqa: E402
import hotcross as H  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="les256")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp32")
    ap.add_argument("--steps", type=int, default=1800)
    ap.add_argument("--spinup", type=int, default=300)
    ap.add_argument("--evals", type=int, default=8)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args()

    p = H.workload_params(args.workload)
    ctx = H.Context(p, H.FP64 if args.precision == "fp64" else H.FP32)
    z, t, s = H.linear_stratification(p)
    ctx.init_case(z, t, s, t_noise=1.0e-4, seed=2024)
    tic = time.time()
    ctx.step(args.spinup)
    ctx.synchronize()
    npatch = ctx.npatch()
    seg, need = AD.plan_segment(ctx, args.steps, args.budget_gb * 1e9)
    print("spin-up %d steps %.1f s; %d source cells; checkpoint every %d steps, %.1f GB of checkpoints and tapes"
          % (args.spinup, time.time() - tic, npatch, seg, need / 1e9), flush=True)

    rng = np.random.default_rng(args.seed)
    truth = 1.0 + 0.4 * (rng.random(npatch) - 0.5)
    truth /= truth.mean()
    probes = AD.plane_probes(p)
    every = AD.samples_every(p)
    ns = args.steps // every
    # synthetic data: the window's probe samples under the true map
    ctx.set_controls(truth, 1.0, p.cs_h)
    ctx.set_probes(probes, every, 1.0, 1.0)
    tic = time.time()
    _, obs, _, _ = ctx.adj_window(args.steps, seg, np.zeros((ns, len(probes), 2)), end_state=False)
    print("synthetic data: one window %.1f s" % (time.time() - tic), flush=True)

    # weights that make w (m/s) and T (K) misfits comparable
    win = AD.Window(ctx, probes, every, args.steps, seg, obs, wgt_w=1.0e4, wgt_t=1.0e4)
    q = torch.ones(npatch, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(1.0, dtype=torch.float64)
    cs = torch.tensor(p.cs_h, dtype=torch.float64)
    opt = torch.optim.LBFGS([q], lr=1.0, max_iter=args.evals, max_eval=args.evals, tolerance_grad=0.0,
                            tolerance_change=0.0, history_size=10, line_search_fn="strong_wolfe")
    hist = []

    def closure():
        opt.zero_grad()
        t0 = time.time()
        loss = win.loss(q, a, cs)
        loss.backward()
        err = float(np.linalg.norm(q.detach().numpy() - truth) / np.linalg.norm(truth - 1.0))
        hist.append({"loss": float(loss.detach()), "map_err": err, "seconds": time.time() - t0})
        print("evaluation %d: loss %.6e, map error %.4f (relative to the true deviation), %.1f s"
              % (len(hist), hist[-1]["loss"], err, hist[-1]["seconds"]), flush=True)
        return loss

    opt.step(closure)
    out = {"workload": args.workload, "precision": args.precision, "steps": args.steps, "seg": seg,
           "hbm_gb": need / 1e9, "npatch": npatch, "evaluations": hist,
           "loss_ratio": hist[-1]["loss"] / hist[0]["loss"], "map_err_final": hist[-1]["map_err"]}
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
