"""The benched native loop (hc_step: MOMENTUM fused with the work velocities
and REFRESH's filter, pass 1 of T, S and C in one launch, REFRESH as a buffer
rotation, the counter-based source noise) against the oracle replaying the
same start state and the same noise draws (hc_source_noise), on the linear
equation of state of the benched workloads.

Grids: les64 (64^3, dx = 1 m, dt = 0.25 s) for 50 steps, and a ragged
72 x 36 x 40 grid whose x and y extents are not multiples of any kernel tile
(32-wide in x; 16, 8 rows in y) and whose depth splits into several plane
chunks, so that every x, y and z seam of the tiled kernels (corrective
MPDATA pass, pass 1, FORCES, VISCOSITY3D) and DCT lengths 64, 72 and 36 meet
an independent result.  les256 (the headline grid) for 8 steps where the
host has the memory.  Bar: 1e-9 relative to each field's largest magnitude
(the whole-run bar of the golden replays); rhop within the density
difference the T and S differences imply under the linear equation of state.

The replays start from rest plus 1e-4 K noise and run long enough for the
vent to set the flow moving: after two steps v and w are still about 1e-8 m/s
and the pressure 1e-7, so FP64 roundoff on the 0.03 m/s crossflow (1e-15)
is 1e-8 of them, and the pressure's nearly constant modes, which the 256^3
Neumann problem determines only to about 1e-11, are 1e-5 of it; by step 8
every field agrees to 3.4e-10 or better."""
import os

import numpy as np
import pytest

import hc_testlib as T
import gpu_common as G

pytestmark = pytest.mark.gpu
H = G.H
RHO_ULPS = 2 * float(np.spacing(1030.0))
COMPARE = ("u", "v", "w", "um", "vm", "wm", "p", "t", "s", "c", "ss", "ss_v", "rhop")


def _params(name):
    if name == "ragged":
        p = H.workload_params("les64")
        p.nx, p.ny, p.nz = 72, 36, 40
        p.zend = p.zstart + p.nz * p.dz
        return p
    return H.workload_params(name)


def _pair(p, seed):
    """A context set up by hc_init_case and an oracle holding the same start
    state (profiles and sponge from the same table, fields downloaded)."""
    z, t, s = H.linear_stratification(p)
    ctx = H.Context(p, H.FP64)
    ctx.init_case(z, t, s, t_noise=1.0e-4, seed=seed)
    o = T.Oracle(T.Params.from_buffer_copy(bytes(p)))
    o.init_profiles(z, t, s)
    for c in ("density", "initialize", "sponge_maker"):
        o.call(c)
    for f in G.UPLOAD:
        o.set(f, ctx.download(f))
    o.set_time(-p.dt, 0)
    ctx._t_bkg, ctx._s_bkg = o.get("t_bkg"), o.get("s_bkg")
    return ctx, o


def _check(ctx, o, tol, label):
    """Every field within tol of its largest magnitude; rhop (linear
    equation of state) within what the T and S differences imply,
    rho_ref (alpha |dT| + beta |dS|), plus four ulps of rho for the
    subtraction of rho_bkg."""
    worst, absd = {}, {}
    p = ctx.prm
    for f in COMPARE:
        a, b = ctx.download(f), o.get(f)
        if f == "rhop":
            err = G.compare_abs(a, b)
            bar = 4 * RHO_ULPS + 1.01 * p.eos_rho_ref * (p.eos_alpha * absd["t"] + p.eos_beta * absd["s"])
            assert err <= bar, (label, f, err, bar)
        else:
            err = G.compare(a, b)
            assert err <= tol, (label, f, err)
            absd[f] = G.compare_abs(a, b)
        worst[f] = err
    print(label, " ".join("%s %.1e" % kv for kv in worst.items()), flush=True)


@pytest.mark.parametrize("name,nsteps", [("ragged", 12), ("les64", 50)])
def test_native_loop_matches_oracle(name, nsteps):
    p = _params(name)
    print("native loop vs oracle:", name, p.nx, p.ny, p.nz, nsteps, "steps", flush=True)
    ctx, o = _pair(p, seed=21)
    try:
        ctx.step(nsteps)
        for m in range(nsteps):
            o.step(m, ctx.source_noise(m))
        _check(ctx, o, 1e-9, name)
    finally:
        ctx.close()


def test_subroutine_path_matches_oracle_on_ragged_grid():
    """The per-subroutine entries (the Fortran drop-in's path) on the ragged
    multi-tile grid, each step against the oracle."""
    p = _params("ragged")
    print("subroutine path vs oracle: ragged, 6 steps", flush=True)
    ctx, o = _pair(p, seed=22)
    try:
        for m in range(6):
            noise = ctx.source_noise(m)
            ctx.driver_step(m, noise)
            o.step(m, noise)
        _check(ctx, o, 1e-10, "ragged subroutine path")
    finally:
        ctx.close()


def test_linear_eos_density_matches_oracle():
    """DENSITY under the linear equation of state (eos = 1) of the benched
    workloads, on perturbed T and S, against the oracle."""
    p = _params("ragged")
    ctx, o = _pair(p, seed=23)
    try:
        rng = np.random.default_rng(4)
        for f, amp in (("t", 0.3), ("s", 0.05)):
            a = o.get(f) + amp * (rng.random(T.shape_of(o.bounds(f))) - 0.5)
            o.set(f, a)
            ctx.upload(f, a)
        o.call("density")
        ctx.density()
        a, b = ctx.download("rhop"), o.get("rhop")
        assert G.compare_abs(a, b) <= RHO_ULPS
        assert np.abs(b).max() > 1e-3  # the perturbation moved rho
    finally:
        ctx.close()


def _host_memory_gib():
    try:
        return os.sysconf("SC_PAGE_SIZE") * os.sysconf("SC_PHYS_PAGES") / 2**30
    except (ValueError, OSError):
        return 0.0


@pytest.mark.timeout(900)
@pytest.mark.skipif(_host_memory_gib() < 48, reason="the 256^3 oracle needs about 10 GB of host memory")
def test_native_loop_matches_oracle_headline_grid():
    """les256 (256^3, dx = 0.5 m, dt = 0.2 s, tidal crossflow, noisy vent):
    eight native steps against the oracle (about 17 s of host time per
    step)."""
    p = _params("les256")
    print("native loop vs oracle: les256, 8 steps", flush=True)
    ctx, o = _pair(p, seed=24)
    try:
        ctx.step(8)
        for m in range(8):
            o.step(m, ctx.source_noise(m))
            print("oracle step", m, "done", flush=True)
        _check(ctx, o, 1e-9, "les256")
    finally:
        ctx.close()
