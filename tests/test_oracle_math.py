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


def _strain_case(field, axis, alpha, dx=0.5, dy=2.0, dz=0.25):
    """Oracle with a single linear velocity field (um, vm or wm = alpha times
    the coordinate along axis, over the field's whole box), every other
    velocity zero and no density gradient.  Unequal spacings, so that a
    spacing taken for another (dx for dy in s12) changes the result; powers of
    two for them and for alpha, so that every sum and difference of velocities
    is exact and the 1e-30 floor under the square roots sees exact zeros."""
    p = T.reference_params(16, 12, 10, dx=dx, dy=dy, dz=dz)
    o = T.Oracle(p)
    b = T.bounds(field, 16, 12, 10)
    idx = np.meshgrid(*[np.arange(l, h + 1) for l, h in b], indexing="ij")
    d = (dx, dy, dz)[axis]
    o.set(field, np.asfortranarray(alpha * idx[axis] * d))
    o.set("rho_bkg", np.full(12, 1027.0))
    return p, o


# (field, axis, S_H^2 / alpha^2, S_V^2 / alpha^2) for velocity = alpha * coordinate:
#   du/dx: s11 = alpha           -> S_H^2 = 2 alpha^2
#   du/dy: s12 = alpha/2         -> S_H^2 = 4 (alpha/2)^2 = alpha^2
#   dv/dx: s12 = alpha/2         -> alpha^2
#   dv/dy: s22 = alpha           -> 2 alpha^2
#   dw/dz: s33 = alpha           -> 2 alpha^2
#   du/dz: s13 = -alpha/2 (k increases downward in the differences of :74) -> S_V^2 = alpha^2
#   dw/dx: s13 = alpha/2         -> S_V^2 = alpha^2
#   dv/dz, dw/dy: s23 likewise
STRAINS = [("um", 0, 2.0, 0.0), ("um", 1, 1.0, 0.0), ("vm", 0, 1.0, 0.0), ("vm", 1, 2.0, 0.0),
           ("wm", 2, 2.0, 0.0), ("um", 2, 0.0, 1.0), ("wm", 0, 0.0, 1.0), ("vm", 2, 0.0, 1.0),
           ("wm", 1, 0.0, 1.0)]


@pytest.mark.parametrize("field,axis,h2,v2", STRAINS, ids=["%s_d%s" % (f, "xyz"[a]) for f, a, _, _ in STRAINS])
def test_smagorinsky_strain_components_closed_form(field, axis, h2, v2):
    """Each velocity gradient alone, in closed form: the horizontal
    coefficient ss = (cs_h sqrt(dx dy))^2 sqrt(S_H^2 + 1e-30) with
    S_H^2 = 2 s11^2 + 2 s22^2 + 4 s12^2 + 2 s33^2, epsh = S_H^2 (ss + axmin)
    (smagorinsky_split_mixing-revised2014.f:20-21,34-45,57-59); the vertical
    ss_v = (cs_v dz)^2 sqrt(S_V^2 + 1e-30) sqrt(1 + eps) with
    S_V^2 = 4 s13^2 + 4 s23^2 and Ri = 0 (:74-114)."""
    alpha = 2.0**-6
    p, o = _strain_case(field, axis, alpha)
    o.call("viscosity3d")
    ss = o.getf("ss")[1:16, 1:12, 1:10]
    ssv = o.getf("ss_v")[1:16, 1:12, 1:10]
    coef_h = p.cs_h**2 * p.dx * p.dy
    coef_v = p.cs_v**2 * p.dz**2
    assert np.allclose(ss, coef_h * np.sqrt(h2 * alpha**2 + 1e-30), rtol=1e-13, atol=0)
    assert np.allclose(o.getf("epsh")[1:16, 1:12, 1:10], h2 * alpha**2 * (ss + p.axmin), rtol=1e-12, atol=1e-300)
    assert np.allclose(ssv, coef_v * np.sqrt(v2 * alpha**2 + 1e-30) * np.sqrt(1 + p.eps), rtol=1e-13, atol=0)
    assert np.all(o.getf("ri")[1:16, 1:12, 1:10] == 0.0)


def test_smagorinsky_pure_strain_and_rotation():
    """u = a y, v = a x (pure shear strain): s12 = a, S_H^2 = 4 a^2.
    u = a y, v = -a x (solid rotation): s12 = 0, so ss is the floor
    cs_h^2 dx dy 1e-15."""
    a = 2.0**-5
    for sign, h2 in ((1.0, 4.0), (-1.0, 0.0)):
        p, o = _strain_case("um", 1, a)
        b = T.bounds("vm", 16, 12, 10)
        I = np.meshgrid(*[np.arange(l, h + 1) for l, h in b], indexing="ij")[0]
        o.set("vm", np.asfortranarray(sign * a * I * p.dx))
        o.call("viscosity3d")
        ss = o.getf("ss")[1:16, 1:12, 1:10]
        coef_h = p.cs_h**2 * p.dx * p.dy
        expect = coef_h * np.sqrt(h2 * a * a + 1e-30)
        assert np.allclose(ss, expect, rtol=1e-12, atol=0), (sign, ss.min(), ss.max(), expect)


def test_smagorinsky_ghosts_follow_bcond():
    """x/y ghosts of ss copy the neighbouring interior cell (bcond = 1) or
    the opposite one (bcond = 0, periodic); z ghosts of ss_v copy k = 1 and
    nz (:124-146)."""
    for bc in (1, 0):
        p = T.reference_params(16, 12, 10)
        p.bcondx = p.bcondy = bc
        o = T.Oracle(p)
        rng = np.random.default_rng(11)
        for f in ("um", "vm", "wm"):
            o.set(f, 0.01 * rng.standard_normal(T.shape_of(o.bounds(f))))
        o.call("viscosity3d")
        ss, ssv = o.getf("ss"), o.getf("ss_v")
        src_lo, src_hi = (1, 16) if bc else (16, 1)
        assert np.array_equal(ss[0, 1:12, 1:10], ss[src_lo, 1:12, 1:10])
        assert np.array_equal(ss[17, 1:12, 1:10], ss[src_hi, 1:12, 1:10])
        src_lo, src_hi = (1, 12) if bc else (12, 1)
        assert np.array_equal(ss[1:16, 0, 1:10], ss[1:16, src_lo, 1:10])
        assert np.array_equal(ss[1:16, 13, 1:10], ss[1:16, src_hi, 1:10])
        assert np.array_equal(ssv[1:16, 1:12, 0], ssv[1:16, 1:12, 1])
        assert np.array_equal(ssv[1:16, 1:12, 11], ssv[1:16, 1:12, 10])


def _linear_eos_params():
    sys_path = os.path.join(T.ROOT, "plumemodel-hotcross_amd")
    import sys
    if sys_path not in sys.path:
        sys.path.insert(0, sys_path)
    import hotcross as H
    return H, T.Params.from_buffer_copy(bytes(H.workload_params("les64")))


def test_linear_eos_closed_form():
    """The linear equation of state (eos = 1) of the benched workloads,
    rho = rho_ref (1 - alpha (T - T0) + beta (S - S0)): rho(T0, S0) = rho_ref,
    d rho/dT = -alpha rho_ref, d rho/dS = beta rho_ref, exactly linear."""
    _, p = _linear_eos_params()
    assert p.eos == 1
    r0 = T.rho_rho(p, p.eos_s0, p.eos_t0)
    assert r0 == p.eos_rho_ref
    for dT in (1.0, -0.5, 3.0):
        d = T.rho_rho(p, p.eos_s0, p.eos_t0 + dT) - r0
        assert abs(d - (-p.eos_alpha * p.eos_rho_ref * dT)) <= 1e-12 * p.eos_rho_ref
    for dS in (1.0, -0.25, 2.0):
        d = T.rho_rho(p, p.eos_s0 + dS, p.eos_t0) - r0
        assert abs(d - p.eos_beta * p.eos_rho_ref * dS) <= 1e-12 * p.eos_rho_ref
    # warmer is lighter, saltier is heavier; UNESCO at the same point is
    # within 0.5 kg/m3 (a wrong reference density or a dropped factor is not)
    assert T.rho_rho(p, p.eos_s0, p.eos_t0 + 1) < r0 < T.rho_rho(p, p.eos_s0 + 1, p.eos_t0)
    q = p.copy(); q.eos = 0
    assert abs(T.rho_rho(q, p.eos_s0, p.eos_t0) - r0) < 0.5


def test_linear_stratification_gives_the_stated_buoyancy_frequency():
    """linear_stratification(N^2 = 1e-6) through INITIALIZE_PROFILES and the
    linear equation of state: the background density rho_bkg(k) (k counts
    down from the lid, dz apart) has N^2 = g/rho0 d rho/d depth = 1e-6 1/s^2
    between every pair of levels 1..nz (levels 0 and nz+1 copy their
    neighbours, initialize_profiles80.f:66-67), and DENSITY's
    rhop = rho - rho_bkg is zero at rest."""
    H, p = _linear_eos_params()
    o = T.Oracle(p)
    z, t, s = H.linear_stratification(H.workload_params("les64"))
    o.init_profiles(z, t, s)
    o.call("density"); o.call("initialize")
    rb = o.get("rho_bkg")
    assert rb[0] == rb[1] and rb[p.nz + 1] == rb[p.nz]
    n2 = p.g / p.rho0 * np.diff(rb[1:p.nz + 1]) / p.dz
    assert np.allclose(n2, 1.0e-6, rtol=1e-8, atol=0), (n2.min(), n2.max())
    assert np.abs(o.getf("rhop")[1:p.nx + 1, 1:p.ny + 1, 1:p.nz + 1]).max() < 1e-10


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
