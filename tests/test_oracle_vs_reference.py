"""Every timestep subroutine of the oracle against the reference's own
Fortran, compiled in place by oracle/build_ref.py (skipped where the
reference sources or amdflang are absent), from a perturbed reference state."""
import os
import shutil

import numpy as np
import pytest

import hc_testlib as T
import refcases as R

pytestmark = pytest.mark.skipif(not (os.path.isdir(T.REF_DIR) and shutil.which("amdflang")),
                                reason="reference sources / amdflang not present")

SEQS = {
    "body_force+forces": (["body_force", "forces"], ["fx", "fy", "fz", "azu", "azv"], 1e-15),
    "pressure": (["body_force", "forces", "pressure"], ["p"], 1e-14),
    "momentum": (["body_force", "forces", "pressure", "momentum_u", "momentum_v", "momentum_w"],
                 ["up", "vp", "wp"], 1e-14),
    "source_terms": (["body_force", "source_terms"], ["q_t", "q_s", "q_c", "isource_map"], 0.0),
    "viscosity3d": (["viscosity3d"], ["ss", "ss_v", "ri", "epsh", "epsv"], 1e-15),
    "mixing": (["body_force", "viscosity3d", "mixing"], ["kxt", "kyt", "kzt", "chi_v", "chi_h"], 1e-15),
    "tracers+density": (["body_force", "viscosity3d", "source_terms", "mixing", "work_velocity", "tracer1",
                         "tracer2", "density", "tracer3"], ["t", "s", "c", "rhop", "ud", "vd", "wd"], 1e-11),
    "refresh": (["body_force", "refresh"], ["u", "um", "v", "vm", "w", "wm"], 0.0),
}


@pytest.mark.parametrize("name", list(SEQS))
@pytest.mark.parametrize("steps", [[0], [0, 1]], ids=["m0", "m1"])
def test_subroutine_matches_reference(name, steps):
    calls, outs, tol = SEQS[name]
    prm = T.reference_params(16, 12, 10)
    before, after, noises = T.run_in_ref(R.worker_sequence, prm, seed=1, steps=steps, calls=calls)
    o = R.oracle_from(prm, before)
    o.set_time(-prm.dt, 0)
    R.oracle_sequence(o, steps, calls, noises)
    for n in outs:
        a, b = o.get(n), after[n]
        m = np.abs(b) < 1e30
        assert np.array_equal(a[~m], b[~m]), n
        err = T.rel_err(a, b, m)
        assert err <= max(tol, 3e-15), (n, err)


@pytest.mark.parametrize("bcondy", [1, 0])
def test_viscosity3d_richardson_cutoff_matches_reference(bcondy):
    """VISCOSITY3D of the reference (its csqrt at
    smagorinsky_split_mixing-revised2014.f:112 promoted to COMPLEX, as PGI
    does) on a quiet flow in the reference's stratification, where the
    Richardson number passes the cutoff Ri/3 > 1 + eps in some cells and not
    in others, so both branches of the complex root are compared."""
    prm = T.reference_params(16, 12, 10, bcondy=bcondy)
    before, after, _ = T.run_in_ref(R.worker_sequence, prm, seed=3, steps=[0], calls=["viscosity3d"],
                                    amp=1e-4)
    ri = after["ri"][1:17, 1:13, 1:11]
    cut = ri / 3.0 > 1.0 + prm.eps
    assert 0 < cut.sum() < cut.size, cut.sum()
    assert np.all(after["ss_v"][1:17, 1:13, 1:11][cut] == 0.0)
    o = R.oracle_from(prm, before)
    o.set_time(-prm.dt, 0)
    R.oracle_sequence(o, [0], ["viscosity3d"], [])
    for n in ("ss", "ss_v", "ri", "epsh", "epsv"):
        a, b = o.get(n), after[n]
        assert T.rel_err(a, b) <= 3e-15, (n, T.rel_err(a, b))


def test_initial_state_matches_reference():
    """INITIALIZE_PROFILES, DENSITY, INITIALIZE and SPONGE_MAKER restated."""
    prm = T.reference_params(16, 12, 10)
    before, _, _ = T.run_in_ref(R.worker_sequence, prm, seed=1, steps=[], calls=[])
    o = T.Oracle(prm)
    z, t, s = T.read_profile_file(R.PROFILE)
    o.init_profiles(z, t, s)
    o.call("initialize")
    o.call("sponge_maker")
    for n in ("t_bkg", "s_bkg", "rho_bkg", "azzmin", "alphau", "alphav", "alphae", "alphaw", "alphaee"):
        assert np.array_equal(o.get(n), before[n]) or T.rel_err(o.get(n), before[n]) < 1e-15, n


@pytest.mark.parametrize("seed", [1, 2])
def test_cfl_values_match_reference(seed):
    """CFL_VALUES: the oracle's Courant numbers against the reference's
    printed ones (format 6e15.7) on velocities whose extremes sit on the
    scanned ranges' first and last cells, with larger values just outside
    every range face."""
    prm = T.reference_params(16, 12, 10)
    st, printed = T.run_in_ref(R.worker_cfl, prm, seed=seed)
    o = T.Oracle(prm)
    for n, a in st.items():
        o.set(n, a)
    assert o.call("cfl_values") == 0
    vals = o.scalars()[17:23]
    assert np.allclose(vals, printed, rtol=1e-6, atol=0), (vals, printed)
    # the planted in-range extremes, not the larger values outside the ranges
    assert np.allclose(vals, [0.45, -0.4] * 3, rtol=1e-12, atol=0), vals
