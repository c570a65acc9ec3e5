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
