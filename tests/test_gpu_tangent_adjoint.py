"""The GPU adjoint against the oracle's tangent-linear model (oracle/
liboracle_cs.so, the oracle compiled for complex-step differentiation:
every real becomes x + i h x', h = 1e-30, Squire & Trapp 1998).

Both are exact derivatives of the same discrete map, the tangent of the
oracle's forward code and the adjoint of the GPU's, so the dot-product
identity <lam, TL d> = <ADJ lam, d> holds to rounding, without the
truncation and kink errors that bound a difference quotient (the bars of the
difference tests are 1e-9 to 1e-4; here 1e-10).  Per routine, on the chains
and at the points of the difference tests (the 16 x 12 x 10 golden state with
velocity noise), and over the 8-step window with five probes sampled every
second step, the window's initial-state adjoint (dL/dx0 of the probe-misfit
loss) against the loss's tangent along a random direction.

The start state's roundoff zeros (v and w of 1e-19 on the walls and the lid)
are set to the zeros they stand for and the direction leaves the exact zeros
alone: |x| has no derivative at 0, where both take 0, and at 1e-19 the sign
of the roundoff, which the two codes need not share, would pick the slope."""
import numpy as np
import pytest

import hc_testlib as T
import gpu_common as G
import adjoint as AD  # on the path gpu_common sets
from test_gpu_adjoint import EVERY, NSTEPS, PROBES, SCALE, SEG, base_case, window_case

pytestmark = pytest.mark.gpu
BAR = 1e-10


def _snap(st):
    out = {}
    for f, a in st.items():
        a = np.array(a, dtype=np.float64, copy=True)
        if a.ndim == 3:
            m = np.abs(a) < 1e30
            if m.any():
                a[m & (np.abs(a) <= 1e-12 * np.abs(a[m]).max())] = 0.0
        out[f] = a
    return out


def _direction(rng, a, scale):
    d = rng.standard_normal(a.shape) * scale
    d[~(np.abs(a) < 1e30) | (a == 0.0)] = 0.0
    return d


def _oracle_from(prm, st):
    o = T.OracleCS(prm)
    for k, v in st.items():
        o.set(k, v)
    return o


def _tangent_dot(o, lam):
    return sum(float(np.sum(l * o.get_tan(k))) for k, l in lam.items())


def tl_chain(name, ora, ins, outs, gfwd, gadj, seed=3, scale=None, vel_noise=0.01, adj_pre=None):
    """ora: the chain's oracle calls; gfwd: the GPU's forward calls, whose
    outputs must be the oracle's; gadj: the adjoint calls, made after gfwd
    or, where the adjoint expects another state, after adj_pre instead
    (TRACER31's adjoint recomputes the tracer from its entry value, so it
    follows the work velocities only; REFRESH is affine and needs none)."""
    print("tangent vs adjoint:", name, flush=True)
    prm, st, noise = base_case(vel_noise=vel_noise)
    st = _snap(st)
    rng = np.random.default_rng(seed)
    sc = dict(SCALE, **(scale or {}))
    d = {k: _direction(rng, st[k], sc[k]) for k in ins}

    def oracle_run(tan):
        o = _oracle_from(prm, st)
        o.set_time(0.0, 0)
        o.begin_step(1)
        o.call("body_force")
        o.source_terms(noise[1])
        for k, x in tan.items():
            o.set_tan(k, x)
        for c in ora:
            c(o)
        return o
    o = oracle_run({})
    lam = {}
    for k in outs:
        l = rng.standard_normal(o.get(k).shape)
        l[~(np.abs(o.get(k)) < 1e30)] = 0.0
        lam[k] = l
    # one tangent per input field, so that a difference names its field
    lhs = {k: _tangent_dot(oracle_run({k: d[k]}), lam) for k in ins}
    def gpu_context(calls):
        ctx = G.context_from_state(prm, st)
        ctx.begin_step(1, prm.dt)
        ctx.body_force()
        ctx.source_terms(noise[1])
        for c in calls:
            c(ctx)
        return ctx
    ctx = gpu_context(gfwd)
    try:  # the linearisation points agree (rhop = rho - rho_bkg to the ulps of rho)
        for k in outs:
            a, b = ctx.download(k), o.get(k)
            if k == "rhop":
                assert G.compare_abs(a, b) <= 8 * float(np.spacing(1030.0)), (name, k, G.compare_abs(a, b))
            else:
                assert G.compare(a, b) <= 1e-12, (name, k, G.compare(a, b))
    finally:
        ctx.close()
    ctx = gpu_context(adj_pre if adj_pre is not None else gfwd)
    try:
        ctx.adj_reset()
        for k in outs:
            ctx.adj_upload(k, lam[k])
        for c in gadj:
            c(ctx)
        rhs = {k: float(np.sum(d[k] * ctx.adj_download(k))) for k in ins}
    finally:
        ctx.close()
    tot_l, tot_r = sum(lhs.values()), sum(rhs.values())
    big = max(abs(x) for x in list(lhs.values()) + list(rhs.values()))
    for k in ins:
        print("  %-5s <lam, TL d> %+.15e  <ADJ lam, d> %+.15e" % (k, lhs[k], rhs[k]), flush=True)
    rel = abs(tot_l - tot_r) / max(abs(tot_l), abs(tot_r))
    print("%s: <lam, TL d> %.15e  <ADJ lam, d> %.15e  rel %.1e" % (name, tot_l, tot_r, rel), flush=True)
    for k in ins:  # each field's part, relative to the largest part
        assert abs(lhs[k] - rhs[k]) <= BAR * big, (name, k, lhs[k], rhs[k])
    assert rel <= BAR, (name, tot_l, tot_r)


VEL = ("u", "v", "w", "um", "vm", "wm", "ss", "ss_v", "rhop")


def test_forces():
    tl_chain("forces", [lambda o: o.call("forces")], VEL, ("fx", "fy", "fz"),
             [lambda c: c.forces()], [lambda c: c.adj_forces()])


def test_forces_pressure():
    tl_chain("forces+pressure", [lambda o: o.call("forces"), lambda o: o.call("pressure")], VEL, ("p",),
             [lambda c: c.forces(), lambda c: c.pressure()], [lambda c: c.adj_pressure(), lambda c: c.adj_forces()])


@pytest.mark.parametrize("comp", ["u", "v", "w"])
def test_forces_pressure_momentum(comp):
    out = comp + "p"
    tl_chain("forces+pressure+momentum_" + comp,
             [lambda o: o.call("forces"), lambda o: o.call("pressure"), lambda o: o.call("momentum_" + comp)],
             VEL, (out,),
             [lambda c: c.forces(), lambda c: c.pressure(), lambda c: getattr(c, "momentum_" + comp)()],
             [lambda c: getattr(c, "adj_momentum_" + comp)(), lambda c: c.adj_pressure(), lambda c: c.adj_forces()])


def test_viscosity():
    """Velocities and rhop together (the tangent has no step to size, so the
    two need not be checked apart), and ss and ss_v themselves: VISCOSITY3D
    overwrites every cell of ss but leaves ss_v on the i and j ghost columns
    as it was, so there the adjoint passes through."""
    tl_chain("viscosity3d", [lambda o: o.call("viscosity3d")], ("um", "vm", "wm", "rhop", "ss", "ss_v"),
             ("ss", "ss_v"),
             [lambda c: c.viscosity3d()], [lambda c: c.adj_viscosity3d()], vel_noise=0.2)


def test_density():
    tl_chain("density", [lambda o: o.call("density")], ("t", "s"), ("rhop",),
             [lambda c: c.density()], [lambda c: c.adj_density()])


def test_refresh():
    tl_chain("refresh", [lambda o: o.call("refresh")], ("u", "up", "um", "v", "vp", "vm", "w", "wp", "wm"),
             ("u", "um", "v", "vm", "w", "wm"), [lambda c: c.refresh()], [lambda c: c.adj_refresh()], adj_pre=[])


@pytest.mark.parametrize("bounds", ["unlimited", "reference"])
def test_tracer(bounds):
    """TRACER31 on T with the work velocities and the vent's ghost values
    (MIXING's diffusivities, which the GPU forms from ss and ss_v inside the
    tracer's first pass, are a call of their own in the oracle); with the
    reference's bounds for T the limiter acts at the cells that
    reach the background minimum, and both codes take the same flags."""
    prm, st, _ = base_case()
    tb = st["t_bkg"]
    cmin, cmax = (-1.0e10, 1.0e10) if bounds == "unlimited" else (tb.min(), 1.0e10)
    dt = prm.dt
    vent = prm.tsource + tb[prm.nz]
    tl_chain("tracer31 (%s bounds)" % bounds,
             [lambda o: o.call("mixing"), lambda o: o.call("work_velocity"), lambda o: o.tracer("t", cmin, cmax, dt),
              lambda o: o.vent("t", vent)],
             ("t", "u", "up", "v", "vp", "w", "wp", "ss", "ss_v"), ("t",),
             [lambda c: c.work_velocity(dt), lambda c: c.tracer31("t", tb, cmin, cmax, dt),
              lambda c: c.vent_ghost("t", vent)],
             [lambda c: c.adj_vent_ghost("t"), lambda c: c.adj_tracer31("t", tb, cmin, cmax, dt),
              lambda c: c.adj_work_velocity(dt)], adj_pre=[lambda c: c.work_velocity(dt)])


def test_window_initial_state():
    """The whole window: dL/dx0 from hc_adj_window (the native forward loop
    with checkpoints every 3 steps, the reverse sweep) against the tangent
    of the loss L = sum over samples of (1e4 (w - w_obs)^2 + (T - T_obs)^2)/2,
    the oracle stepping from the same state with the native loop's noise
    draws.  The observations are a run with the heat flux and crossflow
    changed, so the misfit is of the size a calibration meets."""
    prm, st = window_case()
    st = _snap(st)
    prm.cs_h = prm.cs_v = 0.2  # the Smagorinsky control of the window
    ns = NSTEPS // EVERY
    ctx = G.context_from_state(prm, st)
    try:
        npatch = ctx.npatch()
        noise = [ctx.source_noise(m) for m in range(NSTEPS)]
        ctx.set_probes(PROBES, EVERY, 1.0e4, 1.0)
        ctx.set_controls(1.1 + 0.05 * np.sin(np.arange(npatch)), 0.9, 0.2)
        _, obs, _, _ = ctx.adj_window(NSTEPS, SEG, np.zeros((ns, len(PROBES), 2)), end_state=False)
        ctx.set_controls(None, 1.0, 0.2)
        print("window: initial-state adjoint on the GPU", flush=True)
        loss, samp, _, _ = ctx.adj_window(NSTEPS, SEG, obs, end_state=False)
        ins = ("u", "v", "w", "um", "vm", "wm", "t", "s", "ss", "ss_v", "rhop")
        lam0 = {k: ctx.adj_download(k) for k in ins}
    finally:
        ctx.close()
    rng = np.random.default_rng(5)
    d = {k: _direction(rng, st[k], SCALE[k]) for k in ins}
    rhs = sum(float(np.sum(d[k] * lam0[k])) for k in ins)
    print("window: tangent of the oracle over %d steps" % NSTEPS, flush=True)
    o = _oracle_from(prm, st)
    o.set_time(-prm.dt, 0)
    for k in ins:
        o.set_tan(k, d[k])
    bw, bt = o.bounds("w"), o.bounds("t")
    wmax = np.abs(samp[:, :, 0]).max()
    lhs, lo = 0.0, 0.0
    for n in range(NSTEPS):
        o.step(n, noise[n])
        if (n + 1) % EVERY == 0:
            q = (n + 1) // EVERY - 1
            w, t, tw, tt = o.get("w"), o.get("t"), o.get_tan("w"), o.get_tan("t")
            for p, (i, j, k) in enumerate(PROBES):
                iw = (i - bw[0][0], j - bw[1][0], k - bw[2][0])
                it = (i - bt[0][0], j - bt[1][0], k - bt[2][0])
                # the oracle's samples are the GPU's, to the native loop's
                # parity bar (1e-9 of the largest magnitude)
                assert abs(w[iw] - samp[q, p, 0]) <= 1e-9 * wmax, (q, p, w[iw], samp[q, p, 0])
                assert abs(t[it] - samp[q, p, 1]) <= 1e-9 * abs(t[it]), (q, p, t[it], samp[q, p, 1])
                # the misfit weights of the GPU's loss, so that the forward's
                # rounding stays out of the comparison of the derivatives
                dw, dT = samp[q, p, 0] - obs[q, p, 0], samp[q, p, 1] - obs[q, p, 1]
                lo += 0.5 * (1.0e4 * dw * dw + dT * dT)
                lhs += 1.0e4 * dw * tw[iw] + dT * tt[it]
    assert abs(lo - loss) <= 1e-12 * loss, (lo, loss)
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print("window: loss %.10e  dL (tangent) %.15e  (adjoint) %.15e  rel %.1e" % (loss, lhs, rhs, rel), flush=True)
    assert rel <= 1e-9, (lhs, rhs)


def _ragged_window_case(seed=31, vel_noise=0.01):
    """A 72 x 36 x 40 grid with les64's spacing and physics (the ragged grid
    of the native-loop parity test): x and y extents that are no multiple of
    the adjoint tiles (32 x 8 cells; the donor gathers take 256 columns a
    block), 40 planes that split into several chunks per launch.  The start state is
    hc_init_case's plus velocity noise and random viscosities, so that every
    term of the step is active; the forcing ramps are cut as in the 16 x 12 x
    10 window.  Returns the parameters, the profile inputs and the state."""
    H = G.H
    p = H.workload_params("les64")
    p.nx, p.ny, p.nz = 72, 36, 40
    p.zend = p.zstart + p.nz * p.dz
    p.flow_ramp_hrs = 0.0
    p.source_ramp_hrs = 1.0e-3
    p.wramp_hrs = 1.0e-3
    prm = T.Params.from_buffer_copy(bytes(p))
    z, t, s = H.linear_stratification(p)
    o = T.Oracle(prm)
    o.init_profiles(z, t, s)
    for c in ("density", "initialize", "sponge_maker"):
        o.call(c)
    st = {k: o.get(k) for k in ("t_bkg", "s_bkg", "rho_bkg", "azzmin", "alphau", "alphav", "alphae", "alphaw")}
    ctx = H.Context(p, H.FP64)
    try:
        ctx.init_case(z, t, s, t_noise=1.0e-4, seed=seed)
        for f in G.UPLOAD:
            st[f] = ctx.download(f)
    finally:
        ctx.close()
    rng = np.random.default_rng(seed)
    for f in ("u", "v", "w", "um", "vm", "wm", "up", "vp", "wp"):
        a = st[f].copy()
        core = np.abs(a) < 1e30
        a[core] += vel_noise * (rng.random(core.sum()) - 0.5)
        st[f] = a
    for f in ("ss", "ss_v"):
        st[f] = 1e-4 * (1.0 + rng.random(st[f].shape))
    return prm, (z, t, s), st


def test_window_initial_state_multi_tile():
    """The window's initial-state adjoint against the oracle's tangent, as in
    test_window_initial_state, on the ragged multi-tile grid: every x, y and z
    seam of the adjoint's tile kernels (LDS rings flushed by atomics at the
    tile borders and chunk ends), of the donor gathers' plane marches and of
    the pressure adjoint's transforms (lengths 72 and 36) lies inside the
    region the probes' adjoints reach in 8 steps (25 probes 20 m above the
    vent, 4 cells apart, and 3 on the tiles' x seam)."""
    prm, (z, t, s), st = _ragged_window_case()
    st = _snap(st)
    prm.cs_h = prm.cs_v = 0.2
    probes = AD.plane_probes(prm, n=5, spacing=4) + [(32, 18, 20), (33, 17, 21), (64, 9, 20)]
    nsteps, seg, every = 8, 3, 2
    ns = nsteps // every
    print("ragged window: %d x %d x %d, %d steps, %d probes" % (prm.nx, prm.ny, prm.nz, nsteps, len(probes)),
          flush=True)
    ctx = G.context_from_state(prm, st)
    try:
        npatch = ctx.npatch()
        noise = [ctx.source_noise(m) for m in range(nsteps)]
        ctx.set_probes(probes, every, 1.0e4, 1.0)
        ctx.set_controls(1.1 + 0.05 * np.sin(np.arange(npatch)), 0.9, 0.2)
        _, obs, _, _ = ctx.adj_window(nsteps, seg, np.zeros((ns, len(probes), 2)), end_state=False)
        ctx.set_controls(None, 1.0, 0.2)
        loss, samp, _, _ = ctx.adj_window(nsteps, seg, obs, end_state=False)
        ins = ("u", "v", "w", "um", "vm", "wm", "t", "s", "ss", "ss_v", "rhop")
        lam0 = {k: ctx.adj_download(k) for k in ins}
    finally:
        ctx.close()
    rng = np.random.default_rng(7)
    d = {k: _direction(rng, st[k], SCALE[k]) for k in ins}
    rhs = sum(float(np.sum(d[k] * lam0[k])) for k in ins)
    print("ragged window: tangent of the oracle", flush=True)
    o = T.OracleCS(prm)
    o.init_profiles(z, t, s)
    for c in ("density", "initialize", "sponge_maker"):
        o.call(c)
    for k in G.UPLOAD:
        o.set(k, st[k])
    o.set_time(-prm.dt, 0)
    for k in ins:
        o.set_tan(k, d[k])
    bw, bt = o.bounds("w"), o.bounds("t")
    wmax = np.abs(samp[:, :, 0]).max()
    lhs, lo = 0.0, 0.0
    for n in range(nsteps):
        o.step(n, noise[n])
        if (n + 1) % every == 0:
            q = (n + 1) // every - 1
            w, tt_, tw, tt = o.get("w"), o.get("t"), o.get_tan("w"), o.get_tan("t")
            for pi, (i, j, k) in enumerate(probes):
                iw = (i - bw[0][0], j - bw[1][0], k - bw[2][0])
                it = (i - bt[0][0], j - bt[1][0], k - bt[2][0])
                assert abs(w[iw] - samp[q, pi, 0]) <= 1e-9 * wmax, (q, pi, w[iw], samp[q, pi, 0])
                assert abs(tt_[it] - samp[q, pi, 1]) <= 1e-9 * abs(tt_[it]), (q, pi, tt_[it], samp[q, pi, 1])
                dw, dT = samp[q, pi, 0] - obs[q, pi, 0], samp[q, pi, 1] - obs[q, pi, 1]
                lo += 0.5 * (1.0e4 * dw * dw + dT * dT)
                lhs += 1.0e4 * dw * tw[iw] + dT * tt[it]
    assert abs(lo - loss) <= 1e-12 * loss, (lo, loss)
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print("ragged window: loss %.10e  dL (tangent) %.15e  (adjoint) %.15e  rel %.1e" % (loss, lhs, rhs, rel),
          flush=True)
    assert rel <= 1e-9, (lhs, rhs)
