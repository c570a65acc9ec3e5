"""Pins of the oracle that do not need the reference: textbook values,
closed forms and invariants chosen so that a dropped term, a wrong sign or a
transposed index fails one of them."""
import ctypes
import os

import numpy as np
import pytest

import hc_testlib as T


def test_unesco_one_atmosphere_check_values():
    # Fofonoff & Millard (1983) check values of the one-atmosphere density
    p = T.reference_params(16, 12, 10)
    assert abs(T.rho_rho(p, 35.0, 5.0) - 1027.67547) < 5e-6
    assert abs(T.rho_rho(p, 0.0, 5.0) - 999.96675) < 5e-6
    assert abs(T.rho_rho(p, 35.0, 25.0) - 1023.34306) < 5e-6


@pytest.mark.skipif(not os.path.isdir(T.REF_DIR), reason="reference data not present")
def test_eos_against_reference_profile_table():
    # column 4 of site_stratification_linear.inc is RHO_UN(S,T,0)-1000 from
    # Ferret, printed at FP32 resolution
    p = T.reference_params(16, 12, 10)
    path = os.path.join(T.REF_DIR, "site_stratification_linear.inc")
    rows = open(path).read().splitlines()[5:]
    worst = 0.0
    for ln in rows:
        if not ln.strip():
            continue
        t, s, sig = float(ln[15:30]), float(ln[30:45]), float(ln[45:60])
        worst = max(worst, abs(T.rho_rho(p, s, t) - 1000.0 - sig))
    assert worst < 1e-4


def laplacian_neumann(P, dx, dy, dz):
    """7-point Laplacian with zero-gradient ghosts (the discrete operator
    HS3CRT inverts for LBDCND=MBDCND=NBDCND=3 with zero boundary data)."""
    Q = np.pad(P, 1, mode="edge")
    c = Q[1:-1, 1:-1, 1:-1]
    return ((Q[2:, 1:-1, 1:-1] - 2 * c + Q[:-2, 1:-1, 1:-1]) / dx**2 +
            (Q[1:-1, 2:, 1:-1] - 2 * c + Q[1:-1, :-2, 1:-1]) / dy**2 +
            (Q[1:-1, 1:-1, 2:] - 2 * c + Q[1:-1, 1:-1, :-2]) / dz**2)


def hs3crt(o, F, bd=None):
    nx, ny, nz = F.shape
    z = lambda *s: np.zeros(s)
    bd = bd or (z(ny, nz), z(ny, nz), z(nx, nz), z(nx, nz), z(nx, ny), z(nx, ny))
    bd = [np.asfortranarray(b) for b in bd]
    Fa = np.asfortranarray(F.copy())
    o.L.o_hs3crt_c(o.h, *[b.ctypes.data for b in bd], Fa.ctypes.data)
    return Fa


@pytest.mark.parametrize("grid,d", [((16, 12, 10), (1.0, 1.0, 1.0)), ((24, 16, 12), (0.5, 0.5, 0.5))])
def test_poisson_inverts_the_neumann_laplacian(grid, d):
    p = T.reference_params(*grid, dx=d[0], dy=d[1], dz=d[2])
    o = T.Oracle(p)
    rng = np.random.default_rng(0)
    P = rng.standard_normal(grid)
    X = hs3crt(o, laplacian_neumann(P, *d))
    assert np.abs((X - X.mean()) - (P - P.mean())).max() < 1e-10


def test_poisson_singular_mode_is_the_plane_correction():
    """An incompatible right-hand side (total delta != 0) is solved as if
    delta/(ny*nz) were removed from every cell of the plane i = nx, which is
    the rule the HIP solve applies explicitly."""
    grid = (16, 12, 10)
    p = T.reference_params(*grid)
    o = T.Oracle(p)
    rng = np.random.default_rng(1)
    F = rng.standard_normal(grid)
    delta = F.sum()
    assert abs(delta) > 1.0
    X1 = hs3crt(o, F)
    F2 = F.copy()
    F2[-1, :, :] -= delta / (grid[1] * grid[2])
    X2 = hs3crt(o, F2)
    assert np.abs((X1 - X1.mean()) - (X2 - X2.mean())).max() < 1e-10
    # and X2 solves the compatible system
    assert np.abs(laplacian_neumann(X2, 1, 1, 1) - F2).max() < 1e-9


def test_smagorinsky_linear_shear_closed_form():
    """u = alpha*z, no density gradient: S13 = -alpha/2, ss_v = cs_v^2 dz^2
    |alpha| sqrt(1+eps); strong stable stratification (Ri > 3) kills ss_v."""
    p = T.reference_params(16, 12, 10, dz=0.5)
    o = T.Oracle(p)
    alpha = 0.01
    b = T.bounds("um", 16, 12, 10)
    K = np.arange(b[2][0], b[2][1] + 1)
    um = np.broadcast_to(alpha * K * p.dz, T.shape_of(b)).copy(order="F")
    o.set("um", um)
    o.set("rho_bkg", np.full(12, 1027.0))
    o.call("viscosity3d")
    ssv = o.getf("ss_v")[1:16, 1:12, 1:10]
    expect = p.cs_v**2 * p.dz**2 * np.sqrt(alpha**2 + 1e-30) * np.sqrt(1 + p.eps)
    assert np.allclose(ssv, expect, rtol=1e-14)
    assert np.allclose(o.getf("ri")[1:16, 1:12, 1:10], 0.0)
    # N^2 = 4*alpha^2 * 3 -> Ri = 12 > 3
    n2 = 12 * alpha**2
    drho = n2 * p.rho0 / p.g * p.dz  # per level, density increasing downward
    o.set("rho_bkg", 1027.0 + drho * np.arange(12))
    o.call("viscosity3d")
    assert np.all(o.getf("ss_v")[1:16, 1:12, 1:10] == 0.0)
    ri = o.getf("ri")[1:16, 1:12, 1:10]
    assert np.allclose(ri, n2 / alpha**2, rtol=1e-8)


def _tracer_case(nx=16, ny=12, nz=10):
    p = T.reference_params(nx, ny, nz)
    p.bit5k = p.bit2k = 0       # no diffusion
    o = T.Oracle(p)
    return p, o


def test_mpdata_preserves_a_uniform_field_in_uniform_flow():
    p, o = _tracer_case()
    for n, val in (("worku", 0.3 / p.dt), ("workv", -0.2 / p.dt), ("workw", 0.1 / p.dt)):
        o.set(n, np.full(T.shape_of(o.bounds(n)), val))
    o.set("c", np.full(T.shape_of(o.bounds("c")), 2.5))
    o.tracer("c", 0.0, 1e10, p.dt)
    assert np.all(o.getf("c")[1:16, 1:12, 1:10] == 2.5)


def test_mpdata_unit_courant_number_is_an_exact_shift():
    """Courant number 1 in x with periodic x: the upwind step moves the field
    one cell and the antidiffusive velocity (|U|-U^2)... vanishes."""
    p, o = _tracer_case()
    p2 = p.copy(); p2.bcondx = 0
    o = T.Oracle(p2)
    o.set("worku", np.full(T.shape_of(o.bounds("worku")), 1.0 * p.dx / p.dt))
    rng = np.random.default_rng(3)
    c = 1.0 + rng.random(T.shape_of(o.bounds("c")))
    o.set("c", c)
    o.tracer("c", 0.0, 1e10, p.dt)
    cf = T.FArray(c, o.bounds("c"))
    new = o.getf("c")[1:16, 1:12, 1:10]
    shifted = cf[0:15, 1:12, 1:10]
    assert np.abs(new - shifted).max() < 1e-14


def test_mpdata_positivity_of_the_passive_tracer():
    p, o = _tracer_case()
    rng = np.random.default_rng(4)
    for n in ("worku", "workv", "workw"):
        o.set(n, 0.4 * (rng.random(T.shape_of(o.bounds(n))) - 0.5) / p.dt)
    c = np.where(rng.random(T.shape_of(o.bounds("c"))) > 0.8, 1.0, 0.0)
    o.set("c", c)
    o.tracer("c", 0.0, 1e10, p.dt)
    assert o.getf("c")[1:16, 1:12, 1:10].min() >= -1e-15
