"""Plane statistics (hc_stats_*): the device running sums equal a host
accumulation of the downloaded fields bit for bit, the native loop's
automatic samples equal explicit ones, and the FP32 form's time means on the
plane 20 m above the vent agree with the FP64 form's."""
import numpy as np
import pytest

import gpu_common as G

pytestmark = pytest.mark.gpu
H = G.H


def host_sample(ctx, p, k0, k1):
    """The HC_NSTAT values per cell of planes k0..k1, shaped (NSTAT, nk, ny, nx),
    from downloaded fields (u over i=-1.., v over j=-1.., w over k=-1..)."""
    nx, ny = p.nx, p.ny
    u, v, w, t = (ctx.download(n) for n in ("u", "v", "w", "t"))
    ks = np.arange(k0, k1 + 1)
    # (i, j, k) -> array index: u[i+1, j, k], v[i, j+1, k], w[i, j, k+1], t[i, j, k]
    uc = 0.5 * (u[1:nx + 1, 1:ny + 1][:, :, ks] + u[2:nx + 2, 1:ny + 1][:, :, ks])
    vc = 0.5 * (v[1:nx + 1, 1:ny + 1][:, :, ks] + v[1:nx + 1, 2:ny + 2][:, :, ks])
    wk = w[1:nx + 1, 1:ny + 1][:, :, ks + 1]
    wc = 0.5 * (w[1:nx + 1, 1:ny + 1][:, :, ks] + wk)
    tt = t[1:nx + 1, 1:ny + 1][:, :, ks]
    vals = (wk, tt, uc, vc, wc, uc * uc, vc * vc, wc * wc, tt * tt, wc * tt)
    return np.stack([np.transpose(a, (2, 1, 0)) for a in vals])


def les64(precision, seed=2024):
    p = H.workload_params("les64")
    ctx = H.Context(p, precision)
    z, t, s = H.linear_stratification(p)
    ctx.init_case(z, t, s, t_noise=1.0e-4, seed=seed)
    return p, ctx


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_sums_equal_host_accumulation(precision):
    print("stats sums, precision", precision, flush=True)
    p, a = les64(precision)
    _, b = les64(precision)
    try:
        k0, k1, every, nsamp = 40, 46, 3, 4
        a.stats_reset(k0, k1, 0)
        b.stats_reset(k0, k1, every)
        ref = np.zeros((H.NSTAT, k1 - k0 + 1, p.ny, p.nx))
        for _ in range(nsamp):
            a.step(every)
            a.stats_accumulate()
            ref += host_sample(a, p, k0, k1)
        b.step(every * nsamp + every - 1)  # the last partial interval adds nothing
        sa, na = a.stats_get()
        sb, nb = b.stats_get()
        assert na == nb == nsamp
        assert np.array_equal(sa, ref)
        assert np.array_equal(sb, ref)
        st = a.stats()
        assert np.all(st["tke"] >= 0) and np.all(np.isfinite(st["wt_flux"]))
        a.stats_reset(0, 0)
        with pytest.raises(H.HcError):
            a.stats_get()
        with pytest.raises(H.HcError):
            a.stats_reset(0, p.nz + 1, 1)
    finally:
        a.close()
        b.close()


def plane_means(prec, seed, steps):
    """Time-mean w_up and T anomaly (against the initial T) on the plane 20 m
    above the vent, les64 case, 1 Hz samples over `steps` steps."""
    p, ctx = les64(prec, seed)
    try:
        kp = p.nz - int(round(20.0 / p.dz))
        t0 = ctx.download("t")[1:p.nx + 1, 1:p.ny + 1, kp].T.copy()
        ctx.stats_reset(kp, kp, max(1, int(round(1.0 / p.dt))))
        for c in range(4):
            ctx.step(steps // 4)
            ctx.synchronize()
            print("precision", prec, "seed", seed, "steps", (c + 1) * steps // 4, flush=True)
        st = ctx.stats()
        return st["w_up"][0], st["t"][0] - t0
    finally:
        ctx.close()


def profile_errors(a, b):
    """Relative L2 differences of the profiles across the plume (the plane's
    means along y and along x) of w_up and the T anomaly, and of the whole
    plane."""
    rel = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))
    errs = {}
    for q, name in enumerate(("w_up", "t_anom")):
        errs[name + "_x"] = rel(a[q].mean(axis=0), b[q].mean(axis=0))
        errs[name + "_y"] = rel(a[q].mean(axis=1), b[q].mean(axis=1))
        errs[name + "_plane"] = rel(a[q], b[q])
    return errs


# Over the first ten simulated minutes (before the trajectories decorrelate)
# the FP32 form's time-mean profiles are within 1% of the FP64 form's
# (measured on MI355X: 1.8e-4 to 3.1e-4).
TOL_PROFILE = 1.0e-2


def test_fp32_time_means_match_fp64():
    steps = 2400
    a64 = plane_means(H.FP64, 2024, steps)
    a32 = plane_means(H.FP32, 2024, steps)
    errs = profile_errors(a32, a64)
    print("fp32 vs fp64 time means, 10 min:", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        if not k.endswith("_plane"):
            assert v <= TOL_PROFILE, (k, v)


def test_fp32_hour_means_within_fp64_spread():
    """One simulated hour: LES trajectories decorrelate, so two FP64 runs
    whose initial temperature noise (1e-4 K) differs only in its seed already
    differ by several percent in their time means.  The FP32 form's difference
    from the FP64 run must stay within 1.5x that spread, profile by profile
    (measured on MI355X: FP32 3.0e-2 to 4.2e-2, FP64 seed spread 3.7e-2 to
    6.9e-2)."""
    steps = 14400
    a64 = plane_means(H.FP64, 2024, steps)
    b64 = plane_means(H.FP64, 2025, steps)
    a32 = plane_means(H.FP32, 2024, steps)
    e32 = profile_errors(a32, a64)
    spread = profile_errors(b64, a64)
    print("1 h fp32 vs fp64:", {k: "%.2e" % v for k, v in e32.items()}, flush=True)
    print("1 h fp64 seed spread:", {k: "%.2e" % v for k, v in spread.items()}, flush=True)
    for k in e32:
        assert e32[k] <= 1.5 * spread[k], (k, e32[k], spread[k])
