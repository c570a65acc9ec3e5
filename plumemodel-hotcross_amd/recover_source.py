#!/usr/bin/env python3
"""Recovery of the vent's heat-flux map from probe data: the forward + adjoint
configuration at 256^3 (dx = 0.5 m, FP32, 1800 steps, 1 Hz w/T probes on the
plane 20 m above the vent, checkpoints under an HBM budget).

A seeded "true" map (a per-cell factor on the vent heat flux over the 8 x 9
source cells, mean 1 so the total stays 50 MW) produces synthetic probe data;
the map is then fitted from a uniform start by projected gradient descent
(the gradient's mean removed and the map kept at mean 1, so every iterate
keeps the 50 MW total; a safeguarded step length: a trial that lowers the
loss is taken and the step grows, one that does not is dropped and the step
halves), one adjoint window (forward with checkpoints, reverse sweep) per
iteration, or by L-BFGS through the torch harness (adjoint.Window).

usage: python recover_source.py [--workload les256] [--precision fp32]
           [--steps 1800] [--spinup 300] [--iters 20] [--optimizer gd|lbfgs]
           [--checkpointing binomial|fixed|auto] [--budget-gb 200] [--log FILE]
Prints one line per iteration (loss, map error) and a JSON summary; --log
also writes each iteration as a JSON line as it ends.
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


def plan(ctx, steps, budget, kind):
    """(seg, nsnap, bytes, description) of the window's checkpointing."""
    if kind == "auto":
        kind, k, need = AD.plan_window(ctx, steps, budget)
    elif kind == "fixed":
        k, need = AD.plan_segment(ctx, steps, budget)
    else:
        lo = 1
        if ctx.adj_window_binomial_bytes(steps, 1) > budget:
            raise ValueError("not even one snapshot fits the budget")
        hi = steps
        while lo < hi:
            mid = (lo + hi + 1) // 2
            if ctx.adj_window_binomial_bytes(steps, mid) <= budget:
                lo = mid
            else:
                hi = mid - 1
        k, need = lo, ctx.adj_window_binomial_bytes(steps, lo)
    if kind == "fixed":
        return k, None, need, "a checkpoint every %d steps" % k
    nfwd = H.binomial_schedule(steps, k)[1]
    return 1, k, need, "binomial checkpointing, %d snapshot slots, %d forward steps" % (k, nfwd)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="les256")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp32")
    ap.add_argument("--steps", type=int, default=1800)
    ap.add_argument("--spinup", type=int, default=300)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--optimizer", choices=["gd", "lbfgs"], default="gd")
    ap.add_argument("--checkpointing", choices=["binomial", "fixed", "auto"], default="binomial")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--log", default=None)
    args = ap.parse_args()

    p = H.workload_params(args.workload)
    ctx = H.Context(p, H.FP64 if args.precision == "fp64" else H.FP32)
    z, t, s = H.linear_stratification(p)
    ctx.init_case(z, t, s, t_noise=1.0e-4, seed=2024)
    tic = time.time()
    ctx.step(args.spinup)
    ctx.synchronize()
    npatch = ctx.npatch()
    seg, nsnap, need, how = plan(ctx, args.steps, args.budget_gb * 1e9, args.checkpointing)
    print("spin-up %d steps %.1f s; %d source cells; %s, %.1f GB of checkpoints and tapes"
          % (args.spinup, time.time() - tic, npatch, how, need / 1e9), flush=True)

    rng = np.random.default_rng(args.seed)
    truth = 1.0 + 0.4 * (rng.random(npatch) - 0.5)
    truth /= truth.mean()
    probes = AD.plane_probes(p)
    every = AD.samples_every(p)
    ns = args.steps // every
    # synthetic data: the window's probe samples under the true map
    win = AD.Window(ctx, probes, every, args.steps, seg, np.zeros((ns, len(probes), 2)), wgt_w=1.0e4,
                    wgt_t=1.0e4, nsnap=nsnap)
    tic = time.time()
    _, obs, _, _ = win.evaluate(truth, 1.0, p.cs_h)
    print("synthetic data: one window %.1f s" % (time.time() - tic), flush=True)
    win.obs = obs
    hist = []
    log = open(args.log, "w") if args.log else None

    def record(q, loss, g, t0, step=None, **extra):
        err = float(np.linalg.norm(q - truth) / np.linalg.norm(truth - 1.0))
        h = {"iteration": len(hist), "loss": float(loss), "map_err": err, "grad_norm": float(np.linalg.norm(g)),
             "step": step, "seconds": time.time() - t0}
        h.update(extra)
        hist.append(h)
        print("iteration %d: loss %.6e, map error %.4f (relative to the true deviation), |grad| %.3e, %.1f s"
              % (h["iteration"], h["loss"], err, h["grad_norm"], h["seconds"]), flush=True)
        if log:
            log.write(json.dumps(h) + "\n")
            log.flush()

    if args.optimizer == "gd":
        # projected gradient descent with a safeguarded step: each iteration
        # evaluates one trial map q - eta g / max|g| (the largest cell moves
        # by eta), kept within [0.2, 2] and at mean 1 (total 50 MW); a trial
        # that lowers the loss is taken and eta grows by 1.5, one that does
        # not is dropped and eta halves.  Over 1800 steps the plume's
        # trajectory is chaotic in the map, so unguarded steps (Barzilai-
        # Borwein) overshoot until the flow blows up.
        def project(x):
            x = np.clip(x, 0.2, 2.0)
            return x - x.mean() + 1.0

        eta = 0.05
        q_best = np.ones(npatch)
        t0 = time.time()
        loss_best, _, gq, _ = win.evaluate(q_best, 1.0, p.cs_h)
        g_best = gq - gq.mean()  # the gradient on maps of fixed total
        record(q_best, loss_best, g_best, t0, eta)
        while len(hist) <= args.iters:
            trial = project(q_best - eta * g_best / max(np.abs(g_best).max(), 1e-300))
            t0 = time.time()
            loss, _, gq, _ = win.evaluate(trial, 1.0, p.cs_h)
            g = gq - gq.mean()
            accepted = bool(np.isfinite(loss) and loss < loss_best)
            if accepted:
                q_best, loss_best, g_best = trial, loss, g
                eta *= 1.5
            else:
                eta *= 0.5
            record(trial, loss, g, t0, eta, accepted=accepted, best_loss=float(loss_best),
                   best_map_err=float(np.linalg.norm(q_best - truth) / np.linalg.norm(truth - 1.0)))
    else:
        q = torch.ones(npatch, dtype=torch.float64, requires_grad=True)
        a = torch.tensor(1.0, dtype=torch.float64)
        cs = torch.tensor(p.cs_h, dtype=torch.float64)
        opt = torch.optim.LBFGS([q], lr=1.0, max_iter=args.iters, max_eval=args.iters, tolerance_grad=0.0,
                                tolerance_change=0.0, history_size=10, line_search_fn="strong_wolfe")

        def closure():
            opt.zero_grad()
            t0 = time.time()
            loss = win.loss(q, a, cs)
            loss.backward()
            record(q.detach().numpy().copy(), loss.detach(), q.grad.numpy(), t0)
            return loss

        opt.step(closure)
    best = min(hist, key=lambda h: h["loss"] if np.isfinite(h["loss"]) else np.inf)
    out = {"workload": args.workload, "precision": args.precision, "steps": args.steps, "checkpointing": how,
           "hbm_gb": need / 1e9, "npatch": npatch, "optimizer": args.optimizer, "iterations": hist,
           "loss_ratio": best["loss"] / hist[0]["loss"], "map_err_best": best["map_err"]}
    print(json.dumps(out))
    if log:
        log.close()
    ctx.close()


if __name__ == "__main__":
    main()
