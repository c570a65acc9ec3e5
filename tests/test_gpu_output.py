"""Output writers on the device state: the snapshot and the y-slab records
read back by an independent netCDF reader, FP32 values equal to the FP64
fields rounded once, w stored as -w with the box edges filled; the CFL
bailout writes the emergency snapshot before it raises."""
import os
import sys

import numpy as np
import pytest
from scipy.io import netcdf_file

import hc_testlib as T
import gpu_common as G
from test_oracle_golden import load_case

sys.path.insert(0, os.path.join(T.ROOT, "plumemodel-hotcross_amd"))
import output  # noqa: E402

pytestmark = pytest.mark.gpu
H = G.H
GOLDEN = os.path.join(T.ROOT, "tests", "golden", "ref_16x12x10_50.npz")


def expected(ctx, name):
    """What the writers store: real(field, kind=4), w as -w with the twelve
    edges of its box at -1e34."""
    a = ctx.download(name)
    if name == "w":
        a = -a
        e = [np.isin(np.arange(a.shape[d]), (0, a.shape[d] - 1)) for d in range(3)]
        m = (e[0][:, None, None] & e[1][None, :, None]) | (e[0][:, None, None] & e[2][None, None, :]) | \
            (e[1][None, :, None] & e[2][None, None, :])
        a = a.astype(np.float32)
        a[m] = np.float32(-1e34)
        return a
    return a.astype(np.float32)


def test_snapshot_and_slab_records(tmp_path):
    prm, before, _, noise, _ = load_case(GOLDEN)
    print("output writers", flush=True)
    ctx = G.context_from_state(prm, before)
    try:
        for m in range(3):
            ctx.driver_step(m, noise[m])
        path = str(tmp_path / "snap.nc")
        output.snapshot(ctx, path, title="golden 16x12x10")
        f = netcdf_file(path, "r", mmap=False)
        try:
            assert f.dimensions["ux"] == prm.nx + 3 and f.dimensions["wz"] == prm.nz + 3
            for var, field in (("u", "u"), ("v", "v"), ("w", "w"), ("p", "p"), ("temp", "t"), ("s", "s")):
                got = f.variables[var][0]
                assert np.array_equal(got, expected(ctx, field).transpose(2, 1, 0)), var
            assert np.array_equal(f.variables["c"][:], expected(ctx, "c").transpose(2, 1, 0))
            time, _, _ = ctx.clock()
            assert f.variables["time"][0] == time
            sx = f.variables["sx"][:]
            assert abs(sx[1] - sx[0] - prm.dx) < 1e-6 and abs(sx[0] - (-(prm.nx / 3.0) * prm.dx - 0.5 * prm.dx)) < 1e-5
        finally:
            f.close()
        # y-slab records, two of them
        spath = str(tmp_path / "slab.nc")
        nys, nye = 4, 7
        with output.SlabWriter(ctx, spath, nys, nye) as sw:
            sw.append()
            ctx.driver_step(3, noise[3])
            sw.append()
        f = netcdf_file(spath, "r", mmap=False)
        try:
            assert f.variables["u"].shape == (2, prm.nz + 2, nye - nys + 1, prm.nx + 3)
            assert f.variables["v"].shape == (2, prm.nz + 2, nye - nys + 2, prm.nx + 2)
            w = expected(ctx, "w")[:, nys:nye + 1, :]
            assert np.array_equal(f.variables["w"][1], w.transpose(2, 1, 0))
            v = expected(ctx, "v")[:, nys:nye + 2, :]   # v over j = nys-1..nye (array index j+1)
            assert np.array_equal(f.variables["v"][1], v.transpose(2, 1, 0))
        finally:
            f.close()
    finally:
        ctx.close()


def test_cfl_bailout_writes_emergency_snapshot(tmp_path):
    prm, before, _, _, _ = load_case(GOLDEN)
    print("emergency snapshot", flush=True)
    state = dict(before)
    # a uniform crossflow at 0.8 Courant in the leapfrog's three levels (a
    # divergence-free flow, which the pressure projection keeps)
    for name in ("u", "um", "up"):
        u = state[name].copy()
        u[u != -1.0e34] = 0.8 * prm.dx / prm.dt
        state[name] = u
    ctx = G.context_from_state(prm, state)
    try:
        path = str(tmp_path / "emergency.nc")
        with pytest.raises(H.HcError) as e:
            output.run(ctx, 20, emergency_path=path)
        assert e.value.code == H.HC_ECFL
        f = netcdf_file(path, "r", mmap=False)
        try:
            assert f.message.startswith(b"EMERGENCY BAILOUT AT TIME STEP = 0")
            assert np.array_equal(f.variables["w"][0], expected(ctx, "w").transpose(2, 1, 0))
        finally:
            f.close()
    finally:
        ctx.close()


DIAG = ("ri", "epsh", "epsv", "chi_v", "chi_h")


@pytest.mark.parametrize("path", ["subroutines", "native"])
def test_diagnostics_match_oracle(path):
    """hc_set_diagnostics: Ri, epsh, epsv (VISCOSITY3D) and chi_v, chi_h
    (MIXING) of the last step against the oracle's after the same three
    steps of the golden run (the oracle's are pinned against the reference
    compiled in place), through the per-subroutine entries and through the
    native loop (the oracle replaying its noise draws); the state is the one
    a run without diagnostics reaches, bit for bit."""
    prm, before, _, noise, _ = load_case(GOLDEN)
    print("diagnostics,", path, flush=True)
    o = T.Oracle(prm)
    for k, v in before.items():
        o.set(k, v)
    o.set_time(-prm.dt, 0)
    ctx = G.context_from_state(prm, before)
    plain = G.context_from_state(prm, before)
    try:
        ctx.set_diagnostics(True)
        for m in range(3):
            nz = noise[m] if path == "subroutines" else ctx.source_noise(m)
            o.step(m, nz)
            if path == "subroutines":
                ctx.driver_step(m, nz)
                plain.driver_step(m, nz)
        if path == "native":
            ctx.step(3)
            plain.step(3)
        for n in DIAG:
            a, b = ctx.diagnostic(n), o.get(n)
            assert np.abs(b).max() > 0, n
            err = float(np.abs(a - b).max() / np.abs(b).max())
            # rho' enters Ri as rho' + rho_bkg: the oracle's rhop differs
            # from the device's by its s**1.5 rounding (a few ulps of 1030)
            assert err <= (1e-8 if n == "ri" else 1e-10), (n, err)
        for k in ("u", "v", "w", "t", "s", "ss", "ss_v", "rhop"):
            assert np.array_equal(ctx.download(k), plain.download(k)), k
    finally:
        ctx.close()
        plain.close()


def test_full_block_and_box_writers(tmp_path):
    """PUT_full_block and PUT_full_block2 files from the device state, read
    back: the diagnostics, rrho and the slab/box extents as the device holds
    them, rounded once to FP32."""
    prm, before, _, noise, _ = load_case(GOLDEN)
    print("full-block writers", flush=True)
    ctx = G.context_from_state(prm, before)
    try:
        ctx.set_diagnostics(True)
        for m in range(2):
            ctx.driver_step(m, noise[m])
        nys, nye = 3, 8
        path = str(tmp_path / "block.nc")
        with output.FullBlockWriter(ctx, path, nys, nye) as w:
            w.append()
        f = netcdf_file(path, "r", mmap=False)
        try:
            for n, d in (("epsh", "epsh"), ("chiv", "chi_v"), ("Ri", "ri")):
                want = ctx.diagnostic(d)[:, nys:nye + 1, :].astype(np.float32).transpose(2, 1, 0)
                assert np.array_equal(f.variables[n][0], want), n
            want = expected(ctx, "t")[:, nys:nye + 1, :].transpose(2, 1, 0)
            assert np.array_equal(f.variables["temp"][0], want)
            rr = f.variables["rrho"][0]
            assert np.all((rr[1:-1, 1:-1, 1:-1] > 1000.0))
        finally:
            f.close()
        box = ((2, 10), (1, 5), (3, 9))
        path = str(tmp_path / "box.nc")
        with output.BoxWriter(ctx, path, box) as w:
            w.append()
        f = netcdf_file(path, "r", mmap=False)
        try:
            (xs, xe), (ys, ye), (zs, ze) = box
            want = expected(ctx, "w")[xs:xe + 1, ys:ye + 1, zs:ze + 2].transpose(2, 1, 0)  # w bounds start at -1
            assert np.array_equal(f.variables["w"][0], want)
            want = ctx.diagnostic("epsv")[xs:xe + 1, ys:ye + 1, zs:ze + 1].astype(np.float32).transpose(2, 1, 0)
            assert np.array_equal(f.variables["epsv"][0], want)
        finally:
            f.close()
    finally:
        ctx.close()
