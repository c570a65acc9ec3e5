"""Properties that hold at the headline size (256^3, FP64), where the oracle
would take hours: the projected velocity is discretely divergence-free off the
i = nx plane (where the pressure solve's compatibility correction acts), the
pressure has zero mean, the state stays finite and the native loop
(hc_step) reproduces the per-subroutine path bit for bit."""
import numpy as np
import pytest

import gpu_common as G

pytestmark = pytest.mark.gpu
H = G.H


def divergence(u, v, w, p):
    nx, ny, nz = p.nx, p.ny, p.nz
    # arrays are over the u/v/w bounds; interior cells i=1..nx-1, j=1..ny, k=1..nz
    U = u[1:, 1:ny + 1, 1:nz + 1]           # i = 0..nx+1 -> use i=0..nx
    du = (U[1:nx, :, :] - U[0:nx - 1, :, :]) / p.dx
    V = v[1:nx, :, 1:nz + 1]                 # i = 1..nx-1, j = -1..ny+1
    dv = (V[:, 2:ny + 2, :] - V[:, 1:ny + 1, :]) / p.dy
    W = w[1:nx, 1:ny + 1, :]                 # k = -1..nz+1
    dw = (W[:, :, 2:nz + 2] - W[:, :, 1:nz + 1]) / p.dz
    return du + dv + dw


def test_headline_size_projection_and_finiteness():
    p = H.workload_params("les256")
    print("les256 native loop, 4 steps", flush=True)
    ctx = H.Context(p, H.FP64)
    z, t, s = H.linear_stratification(p)
    ctx.init_case(z, t, s, t_noise=1.0e-4, seed=3)
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


def _grid(name):
    p = H.workload_params("les64")
    if name == "ragged":
        # tiles of 32 x 8 cells and 32-plane chunks left partly empty
        p.nx, p.ny, p.nz = 72, 36, 40
        p.zend = p.zstart + p.nz * p.dz
    elif name == "no_c":
        p.toggle_c = 0  # two transported tracers
    return p


# the native loop runs pass 1 of T, S and C in one launch and fuses the work
# velocities and REFRESH's filter into MOMENTUM; the subroutine path runs
# every entry on its own: the two agree bit for bit
@pytest.mark.parametrize("grid,prec", [("les64", "fp64"), ("les64", "fp32"), ("ragged", "fp64"),
                                       ("no_c", "fp32")])
def test_native_loop_equals_subroutine_path(grid, prec):
    p = _grid(grid)
    print("native loop vs subroutine path:", grid, prec, flush=True)
    pr = H.FP64 if prec == "fp64" else H.FP32
    z, t, s = H.linear_stratification(p)
    a = H.Context(p, pr)
    a.init_case(z, t, s, t_noise=1.0e-4, seed=5)
    b = H.Context(p, pr)
    b.init_case(z, t, s, t_noise=1.0e-4, seed=5)
    b._t_bkg = None
    # the subroutine path needs the host profiles: rebuild them as hc_init_case does
    import hc_testlib as T
    o = T.Oracle(T.Params.from_buffer_copy(bytes(p)))
    o.init_profiles(z, t, s)
    b._t_bkg, b._s_bkg = o.get("t_bkg"), o.get("s_bkg")
    n = 12
    a.step(n)
    for m in range(n):
        b.driver_step(m, a.source_noise(m))
    for f in ("u", "v", "w", "p", "t", "s", "c", "ss_v", "rhop"):
        assert np.array_equal(a.download(f), b.download(f)), f
    a.close()
    b.close()


def test_native_loop_fused_equals_unfused(monkeypatch):
    """HC_NO_FUSE=1 (read when a context is created) runs the native loop
    with separate work-velocity, REFRESH-copy and per-tracer pass-1 kernels;
    both forms give the same fields bit for bit."""
    p = H.workload_params("les64")
    z, t, s = H.linear_stratification(p)
    ctxs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("HC_NO_FUSE", flag)
        c = H.Context(p, H.FP64)
        c.init_case(z, t, s, t_noise=1.0e-4, seed=9)
        ctxs.append(c)
    print("native loop, fused and unfused, 12 steps", flush=True)
    for c in ctxs:
        c.step(12)
    for f in ("u", "v", "w", "p", "t", "s", "c", "ss_v", "rhop"):
        assert np.array_equal(ctxs[0].download(f), ctxs[1].download(f)), f
    for c in ctxs:
        c.close()
