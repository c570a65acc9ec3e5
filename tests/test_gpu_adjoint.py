"""The discrete adjoint on the GPU (FP64), through the C-ABI.

Per-routine tests: for a forward chain F of subroutine calls and the adjoint
chain in reverse order, <lam, (F(x+eps d) - F(x-eps d)) / 2eps> must equal
<d, F'(x)^T lam> for a random direction d over the chain's inputs and a random
adjoint lam over its outputs.  FORCES is quadratic and PRESSURE, MOMENTUM,
REFRESH, the work velocities and the linear equation of state are affine, so
their central differences are exact up to rounding (bar 1e-9); VISCOSITY3D
and TRACER31 are smooth at the random points away from the limiter's
thresholds and |x| kinks (bar 1e-6).

Window tests: hc_adj_window's gradients of the probe-misfit loss with
respect to the controls (heat-flux map, crossflow amplitude, Smagorinsky
constant) against central differences of the loss (bar 1e-4, the stated
target), and the adjoint of the window's initial state against a central
difference of the loss along a random state direction."""
import glob
import os

import numpy as np
import pytest

import hc_testlib as T
import gpu_common as G
from test_oracle_golden import load_case

pytestmark = pytest.mark.gpu

GOLDEN = [p for p in sorted(glob.glob(os.path.join(T.ROOT, "tests", "golden", "*.npz"))) if "16x12x10_50" in p][0]
STATE = ("u", "v", "w", "um", "vm", "wm", "ss", "ss_v", "rhop", "t", "s")


def base_case(seed=11, linear_eos=False, vel_noise=0.01):
    prm, before, _, noise, _ = load_case(GOLDEN)
    prm = prm.copy()
    if linear_eos:
        prm.eos = 1
    rng = np.random.default_rng(seed)
    st = dict(before)
    for f in ("u", "v", "w", "um", "vm", "wm", "up", "vp", "wp"):
        a = st[f].copy()
        core = np.abs(a) < 1e30
        a[core] += vel_noise * (rng.random(core.sum()) - 0.5)
        st[f] = a
    for f in ("ss", "ss_v"):
        st[f] = 1e-4 * (1.0 + rng.random(st[f].shape))
    return prm, st, noise


def direction(rng, a, scale):
    d = rng.standard_normal(a.shape) * scale
    d[~(np.abs(a) < 1e30)] = 0.0
    return d


SCALE = {"u": 1e-3, "v": 1e-3, "w": 1e-3, "um": 1e-3, "vm": 1e-3, "wm": 1e-3, "up": 1e-3, "vp": 1e-3, "wp": 1e-3,
         "ss": 1e-5, "ss_v": 1e-5, "rhop": 1e-3, "t": 1e-3, "s": 1e-3, "p": 1e-3}


def run_chain(prm, st, noise, pre, calls, outs):
    ctx = G.context_from_state(prm, st)
    try:
        ctx.begin_step(1, prm.dt)
        ctx.body_force()
        ctx.source_terms(noise[1])
        for c in pre:
            c(ctx)
        for c in calls:
            c(ctx)
        return {k: ctx.download(k) for k in outs}
    finally:
        ctx.close()


def chain_dot(name, fwd, ins, outs, adj, eps, tol, adj_pre=None, pre=(), seed=3, linear_eos=False, scale=None,
              vel_noise=0.01):
    print("adjoint chain", name, flush=True)
    prm, st, noise = base_case(linear_eos=linear_eos, vel_noise=vel_noise)
    rng = np.random.default_rng(seed)
    sc = dict(SCALE, **(scale or {}))
    d = {k: direction(rng, st[k], sc[k]) for k in ins}
    xp = dict(st)
    xm = dict(st)
    for k in ins:
        xp[k] = st[k] + eps * d[k]
        xm[k] = st[k] - eps * d[k]
    fp = run_chain(prm, xp, noise, pre, fwd, outs)
    fm = run_chain(prm, xm, noise, pre, fwd, outs)
    lam = {}
    lhs = 0.0
    for k in outs:
        l = rng.standard_normal(fp[k].shape)
        l[~((np.abs(fp[k]) < 1e30) & (np.abs(fm[k]) < 1e30))] = 0.0
        lam[k] = l
        lhs += float(np.sum(l * (fp[k] - fm[k]) / (2 * eps)))
    ctx = G.context_from_state(prm, st)
    try:
        ctx.begin_step(1, prm.dt)
        ctx.body_force()
        ctx.source_terms(noise[1])
        for c in pre:
            c(ctx)
        for c in (adj_pre if adj_pre is not None else fwd):
            c(ctx)
        ctx.adj_reset()
        for k in outs:
            ctx.adj_upload(k, lam[k])
        for c in adj:
            c(ctx)
        rhs = sum(float(np.sum(d[k] * ctx.adj_download(k))) for k in ins)
    finally:
        ctx.close()
    print(name, "lhs", lhs, "rhs", rhs, flush=True)
    assert abs(lhs - rhs) <= tol * max(abs(lhs), abs(rhs)), (name, lhs, rhs)


def test_adjoint_forces():
    chain_dot("forces", [lambda c: c.forces()], ("u", "v", "w", "um", "vm", "wm", "ss", "ss_v", "rhop"),
              ("fx", "fy", "fz"), [lambda c: c.adj_forces()], 1.0, 1e-9)


def test_adjoint_forces_pressure():
    chain_dot("forces+pressure", [lambda c: c.forces(), lambda c: c.pressure()],
              ("u", "v", "w", "um", "vm", "wm", "ss", "ss_v", "rhop"), ("p",),
              [lambda c: c.adj_pressure(), lambda c: c.adj_forces()], 1.0, 1e-9)


@pytest.mark.parametrize("comp", ["u", "v", "w"])
def test_adjoint_forces_pressure_momentum(comp):
    mom = {"u": (lambda c: c.momentum_u(), lambda c: c.adj_momentum_u(), "up"),
           "v": (lambda c: c.momentum_v(), lambda c: c.adj_momentum_v(), "vp"),
           "w": (lambda c: c.momentum_w(), lambda c: c.adj_momentum_w(), "wp")}[comp]
    chain_dot("forces+pressure+momentum_" + comp, [lambda c: c.forces(), lambda c: c.pressure(), mom[0]],
              ("u", "v", "w", "um", "vm", "wm", "ss", "ss_v", "rhop"), (mom[2],),
              [mom[1], lambda c: c.adj_pressure(), lambda c: c.adj_forces()], 1.0, 1e-9)


@pytest.mark.parametrize("part", ["velocities", "rhop"])
def test_adjoint_viscosity(part):
    # The Richardson number's factor sqrt(max(0, 1 + eps - Ri/3)) has a
    # square-root singularity at Ri = 3, where a finite difference does not
    # converge: the state's velocity noise (0.2 m/s) keeps every cell's shear
    # far from it (Ri << 1 under this weak stratification).  rhop enters as
    # rhop + rho_bkg (~1027): its step (3e-7) keeps the difference's digits.
    # The two parts are checked apart (their sum cancels).
    ins = ("um", "vm", "wm") if part == "velocities" else ("rhop",)
    chain_dot("viscosity3d " + part, [lambda c: c.viscosity3d()], ins, ("ss", "ss_v"),
              [lambda c: c.adj_viscosity3d()], 1e-4, 1e-6, scale={"um": 1e-4, "vm": 1e-4, "wm": 1e-4, "rhop": 3e-3},
              vel_noise=0.2)


def test_adjoint_density():
    chain_dot("density", [lambda c: c.density()], ("t", "s"), ("rhop",), [lambda c: c.adj_density()], 1.0, 1e-6)


def test_adjoint_refresh():
    chain_dot("refresh", [lambda c: c.refresh()], ("u", "up", "um", "v", "vp", "vm", "w", "wp", "wm"),
              ("u", "um", "v", "vm", "w", "wm"), [lambda c: c.adj_refresh()], 1.0, 1e-9, adj_pre=[])


def _tracer_chain(name, bounds):
    prm, st, _ = base_case()
    tb = st["t_bkg"]
    cmin, cmax = bounds(tb)
    dt = prm.dt
    fwd = [lambda c: c.work_velocity(dt), lambda c: c.tracer31("t", tb, cmin, cmax, dt),
           lambda c: c.vent_ghost("t", prm.tsource + tb[prm.nz])]
    adj = [lambda c: c.adj_vent_ghost("t"), lambda c: c.adj_tracer31("t", tb, cmin, cmax, dt),
           lambda c: c.adj_work_velocity(dt)]
    chain_dot(name, fwd, ("t", "u", "up", "v", "vp", "w", "wp", "ss", "ss_v"), ("t",), adj, 1e-4, 1e-6,
              adj_pre=[lambda c: c.work_velocity(dt)])


def test_adjoint_tracer_unlimited():
    _tracer_chain("tracer31 (limiter inactive)", lambda tb: (-1.0e10, 1.0e10))


def test_adjoint_tracer_limited():
    # the reference's bounds for T: the limiter is active at the cells that
    # reach the background minimum; the flags stay fixed within eps
    _tracer_chain("tracer31 (reference bounds)", lambda tb: (tb.min(), 1.0e10))


# ---------------------------------------------------------------------------
def window_case():
    prm, st, _ = base_case(seed=21)
    prm.flow_ramp_hrs = 0.0
    prm.source_ramp_hrs = 1.0e-3
    prm.wramp_hrs = 1.0e-3
    return prm, st


PROBES = [(5, 6, 6), (6, 6, 7), (7, 5, 8), (4, 7, 5), (6, 6, 9)]
NSTEPS, SEG, EVERY = 8, 3, 2


def window_loss(prm, st, obs, qmap=None, flow=1.0, cs=0.2, dstate=None, eps=0.0, grads=False, nsteps=NSTEPS,
                seg=SEG, every=EVERY, nsnap=None, end=None):
    """The window through hc_adj_window (a checkpoint every seg steps) or,
    with nsnap, hc_adj_window_binomial; end: a dict that receives the
    context's fields after the call (the window's end state)."""
    s2 = dict(st)
    if dstate:
        for k, d in dstate.items():
            s2[k] = st[k] + eps * d
    ctx = G.context_from_state(prm, s2)
    try:
        ctx.set_controls(qmap, flow, cs)
        ctx.set_probes(PROBES, every, 1.0e4, 1.0)
        if nsnap is None:
            loss, samp, gq, gs = ctx.adj_window(nsteps, seg, obs)
        else:
            loss, samp, gq, gs = ctx.adj_window_binomial(nsteps, nsnap, obs)
        lam = {k: ctx.adj_download(k) for k in STATE} if grads else None
        if end is not None:
            for k in STATE:
                end[k] = ctx.download(k)
        return loss, samp, gq, gs, lam
    finally:
        ctx.close()


@pytest.fixture(scope="module")
def window_obs():
    prm, st = window_case()
    ctx = G.context_from_state(prm, st)
    npatch = ctx.npatch()
    ctx.close()
    ns = NSTEPS // EVERY
    truth_q = 1.1 + 0.05 * np.sin(np.arange(npatch))
    _, samp, _, _, _ = window_loss(prm, st, np.zeros((ns, len(PROBES), 2)), qmap=truth_q, flow=0.9, cs=0.18)
    return prm, st, samp, npatch


def test_window_control_gradients(window_obs):
    prm, st, obs, npatch = window_obs
    print("window control gradients", flush=True)
    q0 = np.ones(npatch)
    loss, _, gq, gs, _ = window_loss(prm, st, obs, qmap=q0, flow=1.0, cs=0.2)
    assert loss > 0.0
    # crossflow amplitude
    h = 1e-4
    lp = window_loss(prm, st, obs, qmap=q0, flow=1.0 + h, cs=0.2)[0]
    lm = window_loss(prm, st, obs, qmap=q0, flow=1.0 - h, cs=0.2)[0]
    fd = (lp - lm) / (2 * h)
    print("flow_scale: adjoint", gs[0], "fd", fd, flush=True)
    assert abs(gs[0] - fd) <= 1e-4 * abs(fd), (gs[0], fd)
    # Smagorinsky constant (cs_h = cs_v)
    h = 1e-5
    lp = window_loss(prm, st, obs, qmap=q0, flow=1.0, cs=0.2 + h)[0]
    lm = window_loss(prm, st, obs, qmap=q0, flow=1.0, cs=0.2 - h)[0]
    fd = (lp - lm) / (2 * h)
    print("cs: adjoint", gs[1] + gs[2], "fd", fd, flush=True)
    assert abs(gs[1] + gs[2] - fd) <= 1e-4 * abs(fd), (gs[1] + gs[2], fd)
    # heat-flux map: two cells and the uniform direction
    h = 1e-4
    for e in (np.eye(npatch)[0], np.eye(npatch)[npatch // 2], np.ones(npatch)):
        lp = window_loss(prm, st, obs, qmap=q0 + h * e, flow=1.0, cs=0.2)[0]
        lm = window_loss(prm, st, obs, qmap=q0 - h * e, flow=1.0, cs=0.2)[0]
        fd = (lp - lm) / (2 * h)
        ad = float(gq @ e)
        print("qmap: adjoint", ad, "fd", fd, flush=True)
        assert abs(ad - fd) <= 1e-4 * abs(fd), (ad, fd)


def test_window_fused_equals_unfused(window_obs, monkeypatch):
    """The window's forward steps through the native loop's fused path and,
    with HC_NO_FUSE=1, through the separate kernels: the same loss, and the
    same gradients up to the order of the adjoint's atomic additions."""
    prm, st, obs, npatch = window_obs
    print("window, fused and unfused forward", flush=True)
    q0 = np.ones(npatch)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("HC_NO_FUSE", flag)
        out.append(window_loss(prm, st, obs, qmap=q0, flow=1.0, cs=0.2))
    (l0, s0, q0g, g0, _), (l1, s1, q1g, g1, _) = out
    assert l0 == l1 and np.array_equal(s0, s1)
    assert np.allclose(q0g, q1g, rtol=1e-10, atol=1e-14 * np.abs(q0g).max())
    assert np.allclose(g0, g1, rtol=1e-9, atol=0.0)


def test_window_initial_state_adjoint(window_obs):
    prm, st, obs, npatch = window_obs
    print("window initial-state adjoint", flush=True)
    rng = np.random.default_rng(5)
    ins = ("u", "v", "w", "um", "vm", "wm", "t", "s", "ss", "ss_v", "rhop")
    d = {k: direction(rng, st[k], SCALE[k]) for k in ins}
    _, _, _, _, lam = window_loss(prm, st, obs, grads=True)
    rhs = sum(float(np.sum(d[k] * lam[k])) for k in ins)
    eps = 2e-4
    lp = window_loss(prm, st, obs, dstate=d, eps=eps)[0]
    lm = window_loss(prm, st, obs, dstate=d, eps=-eps)[0]
    lhs = (lp - lm) / (2 * eps)
    print("state: adjoint", rhs, "fd", lhs, flush=True)
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("nsteps,seg", [(7, 3), (7, 1), (1, 1), (9, 4), (8, 3)])
def test_window_short_last_segment(nsteps, seg):
    """Windows whose last segment is a single step (nsteps % seg == 1, a
    checkpoint every step, a one-step window): the reverse sweep starts that
    segment from its checkpoint's host state (model time, FORCES' entrance
    accumulators, step counter).  The crossflow turns over in two minutes, so
    the body force of every step differs, and the gradients are checked
    against central differences of the loss."""
    prm, st = window_case()
    prm.flow_omega = 2.0 * np.pi / 120.0
    print("window", nsteps, "steps, checkpoint every", seg, flush=True)
    ctx = G.context_from_state(prm, st)
    npatch = ctx.npatch()
    ctx.close()
    kw = dict(nsteps=nsteps, seg=seg, every=1)
    obs = np.zeros((nsteps, len(PROBES), 2))
    q0 = np.ones(npatch)
    loss, _, gq, gs, _ = window_loss(prm, st, obs, qmap=q0, **kw)
    assert loss > 0.0
    for name, h, ad, args in (
            ("flow_scale", 1e-4, gs[0], lambda e: dict(flow=1.0 + e)),
            ("cs", 1e-5, gs[1] + gs[2], lambda e: dict(cs=0.2 + e)),
            ("qmap[0]", 1e-4, gq[0], lambda e: dict(qmap=q0 + e * np.eye(npatch)[0]))):
        a = dict(qmap=q0)
        a.update(args(h))
        lp = window_loss(prm, st, obs, **a, **kw)[0]
        a = dict(qmap=q0)
        a.update(args(-h))
        lm = window_loss(prm, st, obs, **a, **kw)[0]
        fd = (lp - lm) / (2 * h)
        print(name, "adjoint", ad, "fd", fd, flush=True)
        # a control the probes cannot see yet (one step) gives fd = 0: the
        # adjoint must then be zero to roundoff of the loss
        assert abs(ad - fd) <= 1e-4 * abs(fd) + 1e-12 * loss, (name, ad, fd)


@pytest.mark.parametrize("nsnap", [1, 2, 3, 9])
def test_window_binomial_equals_fixed_interval(window_obs, nsnap):
    """Binomial checkpointing (hc_adj_window_binomial) against the
    fixed-interval window on the same 8 steps: the same forward (loss,
    samples and end state bit for bit) and the same gradients and
    initial-state adjoint up to the order of the adjoint's atomic additions.
    One slot turns every step from X_0 (28 advances), nine hold every
    state (7 advances)."""
    prm, st, obs, npatch = window_obs
    print("binomial window, %d snapshot slots" % nsnap, flush=True)
    q0 = np.ones(npatch)
    e0, e1 = {}, {}
    l0, s0, gq0, gs0, lam0 = window_loss(prm, st, obs, qmap=q0, grads=True, end=e0)
    l1, s1, gq1, gs1, lam1 = window_loss(prm, st, obs, qmap=q0, grads=True, nsnap=nsnap, end=e1)
    assert l0 == l1 and np.array_equal(s0, s1)
    for k in STATE:
        assert np.array_equal(e0[k], e1[k]), k
    assert np.allclose(gq1, gq0, rtol=1e-10, atol=1e-14 * np.abs(gq0).max())
    assert np.allclose(gs1, gs0, rtol=1e-9, atol=1e-14 * np.abs(gs0).max())
    for k in STATE:
        a, b = lam1[k], lam0[k]
        assert np.abs(a - b).max() <= 1e-10 * max(np.abs(b).max(), 1e-300), k
