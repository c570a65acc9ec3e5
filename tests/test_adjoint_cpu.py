"""The hand-written reverse formulas of the adjoint (hc_adj.h) against
central finite differences of the production forward formulas (hc_cell.h),
both built for the host (tests/adjcheck).  For every routine, random state x,
direction d and output adjoint lam:  <lam, (F(x+eps d) - F(x-eps d)) / 2eps>
must equal <d, F'(x)^T lam>.  FORCES is quadratic, so its central difference
is exact up to rounding; the other formulas are smooth at the random points
(no |x| kink, square root or limiter threshold is crossed within eps)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import hc_testlib as T

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "adjcheck", "adjcheck.hip")
LIB = os.path.join(HERE, "adjcheck", "libadjcheck.so")

D = ctypes.POINTER(ctypes.c_double)
NX, NY, NZ = 10, 8, 6


def build():
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(
            os.path.getmtime(SRC), *(os.path.getmtime(os.path.join(T.ROOT, "plumemodel-hotcross_amd", "csrc", f))
                                     for f in ("hc_adj.h", "hc_cell.h", "hc_mpdata.h", "hc_common.h"))):
        subprocess.run(["hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
                        "-Wno-unused-result", "-o", LIB, SRC], check=True, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT)
    return ctypes.CDLL(LIB)


@pytest.fixture(scope="module")
def lib():
    return build()


def ptr(a):
    return a.ctypes.data_as(D)


def parr(arrs):
    return (D * len(arrs))(*[ptr(a) for a in arrs])


def layout(lib):
    out = (ctypes.c_int64 * 4)()
    lib.ax_layout(NX, NY, NZ, out)
    return out[3]


def params():
    p = T.reference_params(NX, NY, NZ, dx=1.0, dy=1.0, dz=1.0, dt=0.25, bit7=1)
    p.bit9 = 1
    p.bcondy = 1
    return p


def step_args(p, leap=2, delta=1):
    fc = 2.0 * p.omega * np.sin(np.pi * p.lat / 180.0)
    fcy = p.bit9 * 2.0 * p.omega * np.cos(np.pi * p.lat / 180.0)
    acc = [(7 - i) * 0.01 / 7.0 * 3 for i in range(8)]
    a = [p.dx, p.dy, p.dz, p.dt, p.dt, leap, delta, 0.031, 0.032, 0.0021, 0.0022, 1.3e-5, -2.1e-6, fc, fcy] + acc
    return np.array(a, dtype=np.float64)


def dot_check(fwd, adj, x, eps, rng, tol, nout, dscale=None):
    """fwd(x) -> list of outputs; adj(x, lams) -> list of input adjoints.
    dscale: per-input scale of the direction (default 1% of the input's mean
    magnitude)."""
    if dscale is None:
        dscale = [(np.abs(a).mean() + 1e-3) * 1e-2 for a in x]
    d = [rng.standard_normal(a.shape) * sc for a, sc in zip(x, dscale)]
    lams = [rng.standard_normal(fwd(x)[q].shape) for q in range(nout)]
    xp = [a + eps * b for a, b in zip(x, d)]
    xm = [a - eps * b for a, b in zip(x, d)]
    fp, fm = fwd(xp), fwd(xm)
    lhs = sum(float(np.sum(lam * (a - b) / (2 * eps))) for lam, a, b in zip(lams, fp, fm))
    g = adj(x, lams)
    rhs = sum(float(np.sum(gi * di)) for gi, di in zip(g, d))
    assert abs(lhs - rhs) <= tol * max(abs(lhs), abs(rhs), 1e-300), (lhs, rhs)
    return lhs, rhs


def aux_arrays(p, rng):
    nx, ny, nz = p.nx, p.ny, p.nz
    azz = 1e-5 + 1e-4 * rng.random(nz + 2)
    au = 0.05 * rng.random((nx + 3) * (ny + 2))
    av = 0.05 * rng.random((nx + 2) * (ny + 3))
    ae = 0.05 * rng.random((nx + 2) * (ny + 2))
    aw = 0.05 * rng.random(nz + 3)
    return [azz, au, av, ae, aw]


def test_forces_adjoint(lib):
    rng = np.random.default_rng(1)
    p = params()
    n = layout(lib)
    sa = step_args(p)
    aux = aux_arrays(p, rng)
    auxp = parr(aux)
    x = [0.05 * rng.standard_normal(n) for _ in range(6)] + [1e-3 * (1 + rng.random(n)) for _ in range(2)] + \
        [1e-3 * rng.standard_normal(n)]

    def fwd(xs):
        out = [np.zeros(n) for _ in range(3)]
        lib.ax_forces_fwd(ctypes.byref(p), ptr(sa), parr(xs), auxp, ptr(out[0]), ptr(out[1]), ptr(out[2]))
        return out

    def adj(xs, lams):
        g = [np.zeros(n) for _ in range(9)]
        gf = ctypes.c_double(0.0)
        lib.ax_forces_adj(ctypes.byref(p), ptr(sa), ptr(np.zeros(6)), parr(xs), auxp, ptr(lams[0]), ptr(lams[1]),
                          ptr(lams[2]), parr(g), ctypes.byref(gf))
        return g

    dot_check(fwd, adj, x, 1e-2, rng, 1e-10, 3)


def test_forces_flow_scale_derivative(lib):
    """d/d(flow_scale) through ubkg_m, vbkg_m, fbx, fby of FORCES."""
    rng = np.random.default_rng(2)
    p = params()
    n = layout(lib)
    aux = aux_arrays(p, rng)
    x = [0.05 * rng.standard_normal(n) for _ in range(6)] + [1e-3 * (1 + rng.random(n)) for _ in range(2)] + \
        [1e-3 * rng.standard_normal(n)]
    sa = step_args(p)
    fs = np.array([0.7, 0.0, 0.3, 0.0, 2e-4, -1e-4])  # d(ubkg_m, ubkg, vbkg_m, vbkg, fbx, fby)
    lams = [rng.standard_normal(n) for _ in range(3)]

    def fwd(sa_):
        out = [np.zeros(n) for _ in range(3)]
        lib.ax_forces_fwd(ctypes.byref(p), ptr(sa_), parr(x), parr(aux), ptr(out[0]), ptr(out[1]), ptr(out[2]))
        return out

    eps = 1e-3
    sp, sm = sa.copy(), sa.copy()
    for q, slot in ((0, 7), (1, 8), (2, 9), (3, 10), (4, 11), (5, 12)):
        sp[slot] += eps * fs[q]
        sm[slot] -= eps * fs[q]
    fp, fm = fwd(sp), fwd(sm)
    lhs = sum(float(np.sum(l * (a - b) / (2 * eps))) for l, a, b in zip(lams, fp, fm))
    g = [np.zeros(n) for _ in range(9)]
    gf = ctypes.c_double(0.0)
    lib.ax_forces_adj(ctypes.byref(p), ptr(sa), ptr(fs), parr(x), parr(aux), ptr(lams[0]), ptr(lams[1]),
                      ptr(lams[2]), parr(g), ctypes.byref(gf))
    assert abs(lhs - gf.value) <= 1e-9 * abs(lhs), (lhs, gf.value)


@pytest.mark.parametrize("bcondy", [1, 0])
def test_viscosity_adjoint(lib, bcondy):
    rng = np.random.default_rng(3 + bcondy)
    p = params()
    p.bcondy = bcondy
    n = layout(lib)
    sa = step_args(p)
    rho_bkg = 1027.0 + 0.005 * np.arange(p.nz + 2)
    x = [0.05 * rng.standard_normal(n) for _ in range(3)] + [2e-4 * rng.standard_normal(n)]

    def fwd(xs):
        out = [np.zeros(n), np.zeros(n)]
        lib.ax_visc_fwd(ctypes.byref(p), ptr(sa), ptr(xs[0]), ptr(xs[1]), ptr(xs[2]), ptr(xs[3]), ptr(rho_bkg),
                        ptr(out[0]), ptr(out[1]))
        return out

    def adj(xs, lams):
        g = [np.zeros(n) for _ in range(4)]
        lib.ax_visc_adj(ctypes.byref(p), ptr(sa), ptr(xs[0]), ptr(xs[1]), ptr(xs[2]), ptr(xs[3]), ptr(rho_bkg),
                        ptr(lams[0]), ptr(lams[1]), ptr(g[0]), ptr(g[1]), ptr(g[2]), ptr(g[3]))
        return g

    # rhop enters the Richardson number as rhop + rho_bkg (~1027): its
    # direction is large enough that the difference keeps its digits
    dot_check(fwd, adj, x, 1e-5, rng, 1e-6, 2, dscale=[5e-4, 5e-4, 5e-4, 1e-2])


def test_pass1_adjoint(lib):
    rng = np.random.default_rng(5)
    p = params()
    n = layout(lib)
    sa = step_args(p)
    azz = 1e-5 + 1e-4 * rng.random(p.nz + 2)
    cbkg = 10.0 - 0.01 * np.arange(p.nz + 2)
    x = [10.0 + 0.1 * rng.standard_normal(n)] + [0.2 * rng.standard_normal(n) for _ in range(3)] + \
        [1e-3 * (1 + rng.random(n)) for _ in range(2)]

    def fwd(xs):
        out = np.zeros(n)
        lib.ax_pass1_fwd(ctypes.byref(p), ptr(sa), parr(xs), ptr(azz), ptr(cbkg), ptr(out))
        return [out]

    def adj(xs, lams):
        g = [np.zeros(n) for _ in range(6)]
        lib.ax_pass1_adj(ctypes.byref(p), ptr(sa), parr(xs), ptr(azz), ptr(cbkg), ptr(lams[0]), parr(g))
        return g

    dot_check(fwd, adj, x, 1e-6, rng, 1e-7, 1)


def test_antidiffusive_velocity_adjoint(lib):
    rng = np.random.default_rng(6)
    n = layout(lib)
    x = [10.0 + rng.standard_normal(n)] + [0.3 * rng.standard_normal(n) for _ in range(3)]

    def fwd(xs):
        out = [np.zeros(n) for _ in range(3)]
        lib.ax_anti_fwd(NX, NY, NZ, *[ptr(a) for a in xs], *[ptr(o) for o in out])
        return out

    def adj(xs, lams):
        g = [np.zeros(n) for _ in range(4)]
        lib.ax_anti_adj(NX, NY, NZ, *[ptr(a) for a in xs], *[ptr(l) for l in lams], *[ptr(a) for a in g])
        return g

    dot_check(fwd, adj, x, 1e-6, rng, 1e-7, 3)


@pytest.mark.parametrize("flagged", [False, True])
def test_limited_donor_adjoint(lib, flagged):
    rng = np.random.default_rng(7 + flagged)
    n = layout(lib)
    fl = (rng.integers(0, 3, n) if flagged else np.zeros(n)).astype(np.uint8)
    x = [10.0 + rng.standard_normal(n)] + [0.3 * rng.standard_normal(n) for _ in range(3)]

    def fwd(xs):
        out = np.zeros(n)
        lib.ax_donor_fwd(NX, NY, NZ, *[ptr(a) for a in xs], fl.ctypes.data_as(ctypes.c_void_p), ptr(out))
        return [out]

    def adj(xs, lams):
        g = [np.zeros(n) for _ in range(4)]
        lib.ax_donor_adj(NX, NY, NZ, *[ptr(a) for a in xs], fl.ctypes.data_as(ctypes.c_void_p), ptr(lams[0]),
                         *[ptr(a) for a in g])
        return g

    # piecewise linear: a large eps is exact unless a velocity changes sign
    dot_check(fwd, adj, x, 1e-3, rng, 1e-11, 1)


@pytest.mark.parametrize("eos", [0, 1])
def test_density_derivatives(lib, eos):
    p = params()
    p.eos = eos
    lib.ax_rho.restype = ctypes.c_double
    for s, t in ((34.6, 2.0), (35.0, 5.0), (30.0, 12.5), (34.9, 0.5)):
        out = np.zeros(2)
        lib.ax_drho(ctypes.byref(p), ctypes.c_double(s), ctypes.c_double(t), ptr(out))
        h = 1e-5
        fdt = (lib.ax_rho(ctypes.byref(p), ctypes.c_double(s), ctypes.c_double(t + h)) -
               lib.ax_rho(ctypes.byref(p), ctypes.c_double(s), ctypes.c_double(t - h))) / (2 * h)
        fds = (lib.ax_rho(ctypes.byref(p), ctypes.c_double(s + h), ctypes.c_double(t)) -
               lib.ax_rho(ctypes.byref(p), ctypes.c_double(s - h), ctypes.c_double(t))) / (2 * h)
        assert abs(out[0] - fdt) <= 1e-6 * abs(out[0]) + 1e-9, (s, t, out[0], fdt)
        assert abs(out[1] - fds) <= 1e-6 * abs(out[1]) + 1e-9, (s, t, out[1], fds)
