"""The snapshot's further variables (write_snapshota.f:152-240: ri,
kxt/kyt/kzt, Az_bkg, alphaee/alphau/alphav) and the full-block files
(put_block.f:47-376, put_block2.f:47-362), written from a stand-in context
that serves the oracle's state after three steps of the 16 x 12 x 10 golden
run, and read back with scipy.io.netcdf_file.  The host-side pieces are
checked against the oracle: MIXING's diffusivities formed from ss and ss_v
against the oracle's own MIXING (bit for bit, as FP32), alphaee against the
oracle's SPONGE_MAKER, rrho against DENSITY; the diagnostics and fields are
the served values, so the test pins the files' layout (dimensions, variable
order, slab and box extents, w as -w with the box's edges filled)."""
import glob
import os
import sys

import numpy as np
import pytest
from scipy.io import netcdf_file

import hc_testlib as T
from test_oracle_golden import load_case

sys.path.insert(0, os.path.join(T.ROOT, "plumemodel-hotcross_amd"))
import hotcross as H  # noqa: E402
import output  # noqa: E402

GOLDEN = sorted(glob.glob(os.path.join(T.ROOT, "tests", "golden", "ref_16x12x10_50.npz")))[0]


class OracleContext:
    """The parts of hotcross.Context the writers use, over an oracle."""

    def __init__(self, o, prm, time, mtime):
        self.o, self.prm, self._clock = o, prm, (time, mtime, mtime + 1)
        self.diagnostics_on = True

    def bounds(self, name):
        return H.field_bounds(name, self.prm.nx, self.prm.ny, self.prm.nz)

    def download(self, name):
        return self.o.get(name)

    def download_f32(self, name, scale=1.0, lo=None, hi=None):
        a = self.o.get(name)
        blo, _ = self.bounds(name)
        lo = lo or blo
        hi = hi or self.bounds(name)[1]
        sl = tuple(slice(l - b, h - b + 1) for l, h, b in zip(lo, hi, blo))
        return (scale * a[sl]).astype(np.float32)

    def clock(self):
        return self._clock

    def profiles(self):
        return tuple(self.o.get(n) for n in ("t_bkg", "s_bkg", "rho_bkg", "azzmin"))

    def sponge(self):
        return tuple(self.o.get(n) for n in ("alphau", "alphav", "alphae", "alphaw"))

    def diagnostic(self, name):
        return self.o.get(name)


@pytest.fixture(scope="module")
def ctx():
    prm, before, _, noise, _ = load_case(GOLDEN)
    o = T.Oracle(prm)
    for k, v in before.items():
        o.set(k, v)
    o.set_time(-prm.dt, 0)
    for m in range(3):
        o.step(m, noise[m])
    return OracleContext(o, prm, 2 * prm.dt, 2)


def _read(path):
    return netcdf_file(path, "r", mmap=False)


def f32(a):
    return np.asarray(a, dtype=np.float32)


def test_diffusivities_and_sponge_masks_follow_the_oracle(ctx):
    o = ctx.o
    kxt, kyt, kzt = output.mixing_diffusivities(ctx, o.get("ss"), o.get("ss_v"))
    for a, n in ((kxt, "kxt"), (kyt, "kyt"), (kzt, "kzt")):
        assert np.array_equal(a, o.get(n)), n
        assert np.abs(a).max() > 0
    aee, au, av, ae = output.sponge_masks(ctx)
    assert np.array_equal(aee, o.get("alphaee"))
    assert np.array_equal(au, o.get("alphau")) and np.array_equal(ae, o.get("alphae"))


def test_snapshot_variables(ctx, tmp_path):
    p = ctx.prm
    path = str(tmp_path / "snap.nc")
    output.snapshot(ctx, path, title="blocks")
    f = _read(path)
    try:
        names = list(f.variables)
        want = ["u", "v", "w", "p", "ss", "ss_v", "ri", "temp", "s", "c", "kxt", "kyt", "kzt", "Az_bkg",
                "alphaee", "alphau", "alphav"]
        assert [n for n in names if n in want] == [n for n in want if n in names]
        assert set(want) - {"c"} <= set(names)
        o = ctx.o
        # Fortran arrays appear with their dimensions reversed
        assert np.array_equal(f.variables["ri"][0], f32(o.get("ri")).transpose(2, 1, 0))
        assert np.array_equal(f.variables["kxt"][:], f32(o.get("kxt")).transpose(2, 1, 0))
        assert np.array_equal(f.variables["kzt"][:], f32(o.get("kzt")).transpose(2, 1, 0))
        assert f.variables["kxt"].shape == (p.nz + 2, p.ny + 2, p.nx + 3)
        assert f.variables["kyt"].shape == (p.nz + 2, p.ny + 3, p.nx + 2)
        assert f.variables["kzt"].shape == (p.nz + 3, p.ny + 2, p.nx + 2)
        assert np.array_equal(f.variables["Az_bkg"][:], f32(o.get("azzmin")))
        assert np.array_equal(f.variables["alphaee"][:], f32(o.get("alphaee")).transpose(2, 1, 0))
        assert np.array_equal(f.variables["alphau"][:], f32(o.get("alphau")).T)
        assert np.array_equal(f.variables["alphav"][:], f32(o.get("alphav")).T)
        assert f.variables["ri"].long_name == b"Richardson Number"
    finally:
        f.close()


def test_snapshot_without_diagnostics_omits_ri(ctx, tmp_path):
    ctx.diagnostics_on = False
    try:
        path = str(tmp_path / "snap2.nc")
        output.snapshot(ctx, path, az_bkg=False)
        f = _read(path)
        try:
            assert "ri" not in f.variables and "Az_bkg" not in f.variables and "kxt" in f.variables
        finally:
            f.close()
    finally:
        ctx.diagnostics_on = True


@pytest.mark.parametrize("nys,nye", [(0, 13), (4, 7)])
def test_full_block(ctx, tmp_path, nys, nye):
    p, o = ctx.prm, ctx.o
    path = str(tmp_path / "block.nc")
    with output.FullBlockWriter(ctx, path, nys, nye) as w:
        w.append()
        w.append()
    f = _read(path)
    try:
        ny = nye - nys + 1
        assert f.dimensions["sy"] == ny and f.dimensions["vy"] == ny + 1
        assert np.array_equal(f.variables["time"][:], [2 * p.dt] * 2)
        for n, d in (("epsh", "epsh"), ("epsv", "epsv"), ("chih", "chi_h"), ("chiv", "chi_v"), ("Ri", "ri")):
            assert np.array_equal(f.variables[n][1], f32(o.get(d)[:, nys:nye + 1, :]).transpose(2, 1, 0)), n
        assert np.array_equal(f.variables["rrho"][0], f32(o.get("rrho")[:, nys:nye + 1, :]).transpose(2, 1, 0))
        assert np.abs(f.variables["chiv"][0]).max() > 0 and np.abs(f.variables["Ri"][0]).max() > 0
        assert np.array_equal(f.variables["alphae"][:], f32(o.get("alphae")[:, nys:nye + 1]).T)
        assert np.array_equal(f.variables["alphau"][:], f32(o.get("alphau")[:, nys:nye + 1]).T)
        assert np.array_equal(f.variables["alphav"][:], f32(o.get("alphav")[:, nys:nye + 2]).T)
        assert np.array_equal(f.variables["Az_bkg"][:], f32(o.get("azzmin")))
        # u over (-1:nx+1, nys:nye), v over (0:nx+1, nys-1:nye)
        assert np.array_equal(f.variables["u"][0], f32(o.get("u")[:, nys:nye + 1, :]).transpose(2, 1, 0))
        assert np.array_equal(f.variables["v"][0], f32(o.get("v")[:, nys:nye + 2, :]).transpose(2, 1, 0))
        sy = f.variables["sy"][:]
        assert np.allclose(sy[0], -(p.ny / 2.0) * p.dy - 0.5 * p.dy + nys * p.dy)
    finally:
        f.close()


def test_box_writer(ctx, tmp_path):
    p, o = ctx.prm, ctx.o
    box = ((3, 9), (2, 6), (1, 8))
    path = str(tmp_path / "box.nc")
    with output.BoxWriter(ctx, path, box) as w:
        w.append()
    f = _read(path)
    try:
        (xs, xe), (ys, ye), (zs, ze) = box
        assert [f.dimensions[d] for d in ("sx", "sy", "sz", "ux", "vy", "wz")] == \
            [xe - xs + 1, ye - ys + 1, ze - zs + 1, xe - xs + 2, ye - ys + 2, ze - zs + 2]
        assert "p" not in f.variables and "ss" not in f.variables
        u = o.get("u")  # bounds start at i = -1
        assert np.array_equal(f.variables["u"][0], f32(u[xs:xe + 2, ys:ye + 1, zs:ze + 1]).transpose(2, 1, 0))
        w_ = o.get("w")  # bounds start at k = -1
        got = f.variables["w"][0].transpose(2, 1, 0)
        want = f32(-w_[xs:xe + 1, ys:ye + 1, zs:ze + 2])
        # the box's cells on the full box's twelve edges are filled; none here
        assert np.array_equal(got, want)
        t = o.get("t")
        assert np.array_equal(f.variables["temp"][0], f32(t[xs:xe + 1, ys:ye + 1, zs:ze + 1]).transpose(2, 1, 0))
        assert np.all(f.variables["c"][0] == output.BoxWriter.FILL)
        assert np.array_equal(f.variables["Ri"][0], f32(o.get("ri")[xs:xe + 1, ys:ye + 1, zs:ze + 1]).transpose(2, 1, 0))
        assert np.allclose(f.variables["sx"][0], -(p.nx / 3.0) * p.dx - 0.5 * p.dx + xs * p.dx)
        assert np.allclose(f.variables["wz"][0], p.zstart - p.dz + zs * p.dz)
    finally:
        f.close()


def test_box_writer_rejects_a_box_outside_the_grid(ctx, tmp_path):
    with pytest.raises(ValueError):
        output.BoxWriter(ctx, str(tmp_path / "x.nc"), ((0, 18), (0, 5), (0, 5)))
