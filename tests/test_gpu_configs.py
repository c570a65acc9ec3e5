"""The configurations beyond the headline grid, at their full sizes.

C3, the precision form at 384 x 384 x 320 (dx = 0.5 m, dt = 0.2 s, the
physics of the 256^3 workload): the FP64 form's projection leaves the
velocity discretely divergence-free off the i = nx plane (where the pressure
solve's compatibility correction acts), the pressure has zero mean and the
state stays finite; the lengths 384 (radix 3) and 320 of the transforms and
the partial tiles of 384 = 12 x 32 columns and 320 planes are exercised.  The
FP32 form after 50 steps from the same state stays within bounds per field
(relative to each field's largest FP64 magnitude) of about three times what
one MI355X measured (u 3.9e-5, v 6.6e-5, w 2.8e-4, p 3.2e-4, T 2.9e-4,
S 3.4e-6, c 1.3e-5, rhop 3.2e-4).

C4, the adjoint configuration (128^3, dx = 1 m, FP64, 400 steps at dt =
0.25 s, a checkpoint every 20 steps, 25 probes 20 m above the vent sampled
at 1 Hz, observations from a truth run with the heat flux x1.1 and the
crossflow x0.9): the window's gradients with respect to the crossflow
amplitude, the Smagorinsky constant and one cell of the heat-flux map
against central differences of the loss.  The differences need a small step
over 400 steps of a developing plume: for the crossflow amplitude one
MI355X measured the adjoint -11.7132995 against central differences
-20.22 (h = 1e-3), -11.629 (1e-4), -11.7131527 (1e-5), -11.7131360 (1e-6),
-11.722 (1e-7): above 1e-5 the perturbation moves limiter flags and |x|
kinks of the MPDATA passes, below 1e-6 roundoff of the 2.7e3 loss takes
over; in between they agree to 1.4e-5.  For the Smagorinsky constant the
adjoint 1.03600501e5 against 1.0475e5 (h = 1e-4), 1.0336e5 (1e-5),
1.1636e5 (1e-6), 1.03600189e5 (1e-7), 1.03600653e5 (1e-8): the steps that
carry a cell across the Richardson cutoff (the square root of
smagorinsky_split_mixing-revised2014.f:112, whose slope is unbounded there)
are off by up to 12%, the smallest agree to 3e-6 and 1.5e-6.  For the
heat-flux cell: 4.4e-5 at h = 1e-4, 2.6e-5 at 1e-6.

C5, the FP32 forward + adjoint at 256^3 (dx = 0.5 m, dt = 0.2 s, an 8 x 9
source map seeded +-20%): the checkpointing plan_window picks for the
configuration's 1800 steps under 200 GB (a fixed interval of 30, which fits)
and under 60 GB (binomial snapshots), and a 30-step window at the full grid
under binomial checkpointing (3 snapshot slots) against the fixed-interval
window (a checkpoint every 10): the same forward, so the loss and the probe
samples agree bit for bit; the gradients agree to the rounding of the
adjoint's atomic sums in FP32 (measured 3e-5 between two evaluations of the
same window at this size, `tests/test_gpu_adjoint_determinism.py`)."""
import numpy as np
import pytest

import gpu_common as G
from test_gpu_fullsize import divergence

pytestmark = pytest.mark.gpu
H = G.H


def _c3_context(prec):
    p = H.workload_params("les384")
    z, t, s = H.linear_stratification(p)
    ctx = H.Context(p, prec)
    ctx.init_case(z, t, s, t_noise=1.0e-4, seed=31)
    return p, ctx


def test_c3_fp64_projection_and_finiteness():
    p, ctx = _c3_context(H.FP64)
    print("les384 FP64, 4 steps", flush=True)
    try:
        ctx.step(4)
        u, v, w = ctx.download("u"), ctx.download("v"), ctx.download("w")
        for a in (u, v, w):
            core = np.abs(a) < 1e30
            assert np.isfinite(a[core]).all()
        d = divergence(u, v, w, p)
        umax = max(np.abs(u[np.abs(u) < 1e30]).max(), 1e-30)
        assert np.abs(d).max() <= 1e-9 * umax / p.dx, (np.abs(d).max(), umax)
        pr = ctx.download("p")[1:p.nx + 1, 1:p.ny + 1, 1:p.nz + 1]
        assert abs(pr.mean()) <= 1e-12 * max(np.abs(pr).max(), 1e-300)
        for f in ("t", "s", "c", "ss", "ss_v", "rhop"):
            assert np.isfinite(ctx.download(f)).all(), f
    finally:
        ctx.close()


C3_BOUNDS = {"u": 1.2e-4, "v": 2e-4, "w": 8.5e-4, "p": 1e-3, "t": 9e-4, "s": 1e-5, "c": 4e-5, "rhop": 1e-3}


def test_c3_fp32_form_within_bounds_of_fp64_after_50_steps():
    print("les384 FP64 and FP32, 50 steps each", flush=True)
    p, a = _c3_context(H.FP64)
    _, b = _c3_context(H.FP32)
    try:
        a.step(50)
        b.step(50)
        worst = {}
        for f, bar in C3_BOUNDS.items():
            x, y = a.download(f), b.download(f)
            m = np.abs(x) < 1e30
            err = float(np.abs(x[m] - y[m]).max() / max(np.abs(x[m]).max(), 1e-300))
            worst[f] = err
            assert err <= bar, (f, err, bar)
        print("les384 FP32 vs FP64 after 50 steps:", " ".join("%s %.2e" % kv for kv in worst.items()), flush=True)
    finally:
        a.close()
        b.close()


def _c4_window(ctx, obs, qmap, flow, cs, nsteps, seg):
    ctx.set_controls(qmap, flow, cs)
    loss, samp, gq, gs = ctx.adj_window(nsteps, seg, obs, end_state=False)
    return loss, samp, gq, gs


def test_c4_adjoint_gradients_match_finite_differences():
    """C4 in full: every evaluation returns the context to the window's start
    (end_state = False), so the differences see the same initial state."""
    import adjoint as AD
    p = H.workload_params("les128")
    p.toggle_c = 0
    z, t, s = H.linear_stratification(p)
    ctx = H.Context(p, H.FP64)
    try:
        ctx.init_case(z, t, s, t_noise=1.0e-4, seed=41)
        ctx.step(8)  # the flow and plume under way
        nsteps, seg = 400, 20
        probes = AD.plane_probes(p)
        every = AD.samples_every(p)
        ctx.set_probes(probes, every, 1.0e4, 1.0)
        npatch = ctx.npatch()
        ns = nsteps // every
        print("C4: truth run (heat flux x1.1, crossflow x0.9), 400 steps", flush=True)
        _, obs, _, _ = _c4_window(ctx, np.zeros((ns, len(probes), 2)), np.full(npatch, 1.1), 0.9, 0.2, nsteps, seg)
        q0 = np.ones(npatch)
        loss, _, gq, gs = _c4_window(ctx, obs, q0, 1.0, 0.2, nsteps, seg)
        print("C4: loss", loss, flush=True)
        assert loss > 0.0 and np.isfinite(gq).all() and np.isfinite(gs).all()
        checks = (("flow_scale", 1e-5, gs[0], lambda e: (q0, 1.0 + e, 0.2)),
                  ("cs", 1e-7, gs[1] + gs[2], lambda e: (q0, 1.0, 0.2 + e)),
                  ("qmap[%d]" % (npatch // 2), 1e-4, gq[npatch // 2],
                   lambda e: (q0 + e * np.eye(npatch)[npatch // 2], 1.0, 0.2)))
        for name, h, ad, args in checks:
            lp = _c4_window(ctx, obs, *args(h), nsteps, seg)[0]
            lm = _c4_window(ctx, obs, *args(-h), nsteps, seg)[0]
            fd = (lp - lm) / (2 * h)
            print("C4 %s: adjoint %.10e fd %.10e rel %.2e" % (name, ad, fd, abs(ad - fd) / abs(fd)), flush=True)
            assert abs(ad - fd) <= 1e-4 * abs(fd), (name, ad, fd)
    finally:
        ctx.close()


def test_c5_checkpointing_plan_and_binomial_window_at_full_size():
    import adjoint as AD
    p = H.workload_params("les256")
    p.toggle_c = 0
    z, t, s = H.linear_stratification(p)
    ctx = H.Context(p, H.FP32)
    try:
        kind, seg, need = AD.plan_window(ctx, 1800, 200e9)
        print("C5 plan under 200 GB:", kind, seg, need / 1e9, "GB", flush=True)
        assert kind == "fixed" and need <= 200e9
        kind2, nsnap, need2 = AD.plan_window(ctx, 1800, 60e9)
        print("C5 plan under 60 GB:", kind2, nsnap, need2 / 1e9, "GB", flush=True)
        assert kind2 == "binomial" and need2 <= 60e9
        ctx.init_case(z, t, s, t_noise=1.0e-4, seed=5)
        ctx.step(4)
        npatch = ctx.npatch()
        rng = np.random.default_rng(5)
        qmap = 1.0 + 0.2 * (2.0 * rng.random(npatch) - 1.0)
        qmap /= qmap.mean()
        ctx.set_controls(qmap, 1.0, 0.2)
        probes = AD.plane_probes(p)
        every = AD.samples_every(p)
        ctx.set_probes(probes, every, 1.0e4, 1.0)
        nsteps = 30
        obs = np.full((nsteps // every, len(probes), 2), 0.01)
        lf, sf, gqf, gsf = ctx.adj_window(nsteps, 10, obs, end_state=False)
        lb, sb, gqb, gsb = ctx.adj_window_binomial(nsteps, 3, obs, end_state=False)
        rel = float(np.max(np.abs(gqb - gqf)) / np.max(np.abs(gqf)))
        print("C5 window: loss", lf, lb, "map rel", rel, "cs", gsf[1] + gsf[2], gsb[1] + gsb[2], flush=True)
        assert np.isfinite(lf) and lf == lb and np.array_equal(sf, sb)
        assert np.all(np.isfinite(gqf)) and np.all(np.isfinite(gqb))
        assert rel <= 5e-4
        assert abs((gsb[1] + gsb[2]) - (gsf[1] + gsf[2])) <= 5e-4 * abs(gsf[1] + gsf[2])
    finally:
        ctx.close()
