"""Parity of the HIP library (through the C-ABI) with the reference's golden
runs and with the oracle.  FP64 plain form: the reference's arithmetic
evaluated in the same order, so the tolerance is the reference-vs-oracle one
(1e-9 relative after a whole run of up to 50 steps).

The one operation whose rounding differs is s**1.5 in the equation of state:
the kernels round it correctly, the host libm pow (the reference's and the
oracle's) misrounds about 1 argument in 1000 by one ulp.  That moves rho
(about 1027 kg/m3) by one ulp, 2.3e-13, which is large next to rhop = rho -
rho_bkg (1e-5 .. 1e-2), so rhop is compared in absolute terms against the
ulp of rho, and the fields that feel rhop through buoyancy get 1e-10 after
a few steps instead of 1e-13 for one call."""
import glob
import os

import numpy as np
import pytest

import hc_testlib as T
import gpu_common as G
from test_oracle_golden import load_case

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(T.ROOT, "tests", "golden", "*.npz")))
RHO_ULPS = 2 * float(np.spacing(1030.0))  # two ulps of the density itself


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_gpu_replays_reference_run(path):
    prm, before, after, noise, n = load_case(path)
    print("case", os.path.basename(path), "steps", n, flush=True)
    ctx = G.context_from_state(prm, before)
    for m in range(n):
        ctx.driver_step(m, noise[m])
    ctx.synchronize()
    for k in T.STATE_FIELDS:
        a = ctx.download(k)
        b = after[k]
        if not np.isfinite(b).any():
            continue  # a run the reference itself does not carry
        err = G.compare(a, b)
        assert err <= 1e-9, (k, err)


def _random_state(prm, seed):
    z = np.load(GOLDEN[0])
    return {k[7:]: z[k] for k in z.files if k.startswith("before_")}


SEQS = {
    "forces": (["body_force", "forces"], ["fx", "fy", "fz"]),
    "pressure": (["body_force", "forces", "pressure"], ["p"]),
    "momentum": (["body_force", "forces", "pressure", "momentum_u", "momentum_v", "momentum_w"],
                 ["up", "vp", "wp"]),
    "viscosity": (["viscosity3d"], ["ss", "ss_v"]),
    "density": (["density"], ["rhop"]),
    "refresh": (["refresh"], ["u", "um", "v", "vm", "w", "wm"]),
}


@pytest.mark.parametrize("name", list(SEQS))
@pytest.mark.parametrize("m", [0, 1])
def test_gpu_routine_matches_oracle(name, m):
    calls, outs = SEQS[name]
    path = [p for p in GOLDEN if "16x12x10_50" in p][0]
    prm, before, _, noise, _ = load_case(path)
    rng = np.random.default_rng(5 + m)
    state = dict(before)
    for f in ("u", "v", "w", "um", "vm", "wm", "up", "vp", "wp"):
        a = state[f].copy()
        core = a != -1.0e34
        a[core] += 0.01 * (rng.random(core.sum()) - 0.5)
        state[f] = a
    for f in ("ss", "ss_v"):
        state[f] = 1e-4 * (1.0 + rng.random(state[f].shape))
    print("routine", name, "step", m, flush=True)
    o = T.Oracle(prm)
    for k, v in state.items():
        o.set(k, v)
    ctx = G.context_from_state(prm, state)
    time = -prm.dt
    for mm in range(m + 1):
        time = time + prm.dt
        o.begin_step(mm)
        ctx.begin_step(mm, time)
        seq = calls if mm == m else ["body_force"]
        for c in seq:
            o.call(c)
            getattr(ctx, c)()
    for f in outs:
        a, b = ctx.download(f), o.get(f)
        if f == "rhop":
            err = G.compare_abs(a, b)
            assert err <= RHO_ULPS, (name, f, err)
        else:
            err = G.compare(a, b)
            assert err <= 1e-13, (name, f, err)


@pytest.mark.parametrize("m", [0, 1, 2])
def test_gpu_tracers_match_oracle(m):
    path = [p for p in GOLDEN if "16x12x10_50" in p][0]
    prm, before, _, noise, _ = load_case(path)
    print("tracers step", m, flush=True)
    o = T.Oracle(prm)
    for k, v in before.items():
        o.set(k, v)
    o.set_time(-prm.dt, 0)
    ctx = G.context_from_state(prm, before)
    for mm in range(m + 1):
        o.step(mm, noise[mm])
        ctx.driver_step(mm, noise[mm])
    for f in ("t", "s", "c", "rhop", "ss", "ss_v", "u", "v", "w", "p"):
        a, b = ctx.download(f), o.get(f)
        if f == "rhop":
            err = G.compare_abs(a, b)
            assert err <= 4 * RHO_ULPS, (f, err)
        else:
            err = G.compare(a, b)
            assert err <= 1e-10, (f, err)


@pytest.mark.parametrize("seed", [1, 2])
def test_gpu_cfl_values_match_oracle(seed):
    """CFL_VALUES bit for bit against the oracle, on velocities with planted
    range extremes; then the two bailouts: a Courant number above 0.75
    (HC_ECFL, as the oracle signals) and a NaN velocity (HC_ENAN, which the
    min/max scan alone would pass over)."""
    import refcases as R
    path = [p for p in GOLDEN if "16x12x10_50" in p][0]
    prm, before, _, _, _ = load_case(path)
    print("cfl seed", seed, flush=True)
    state = R.plant_cfl_extremes(before, prm, seed)
    o = T.Oracle(prm)
    for k, v in state.items():
        o.set(k, v)
    ctx = G.context_from_state(prm, state)
    try:
        assert o.call("cfl_values") == 0
        assert np.array_equal(ctx.cfl_values(), o.scalars()[17:23])
        # bailout: one interior w over 0.75 Courant
        w = state["w"].copy()
        w[3, 4, 5] = 0.8 * prm.dz / prm.dt
        o.set("w", w)
        ctx.upload("w", w)
        assert o.call("cfl_values") == 1
        with pytest.raises(G.H.HcError) as e:
            ctx.cfl_values()
        assert e.value.code == -4
        # NaN: the oracle's min/max pass over it; the library stops
        u = state["u"].copy()
        u[5, 3, 2] = np.nan
        ctx.upload("w", state["w"])
        ctx.upload("u", u)
        with pytest.raises(G.H.HcError) as e:
            ctx.cfl_values()
        assert e.value.code == -5 and "NaN" in str(e.value)
    finally:
        ctx.close()
