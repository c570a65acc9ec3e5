"""Output writers of the timestep's state, in the reference's netCDF layout:
the snapshot of write_snapshota.f (PUT_snapshot_header/_open/PUT_snapshot/
_close, :4-397; also the emergency file of the CFL bailout,
hotcross80-revised2014.f:406-414) and the y-slab record file of put_block1.f
(PUT_full_block1_OPEN/PUT_full_block1, :1-330; the 1 Hz "cone" output of
README.md:26).

Fields leave the device as FP32 (the writers store real(field, kind=4)),
converted there by hc_download_f32; w is stored as -w with the box's twelve
edges set to -1e34 (ww = -w; fill_value, put_block1.f:265-266,
hotcross80-revised2014.f:1499-1528).  Dimension and variable names, units,
long names and the missing_value attribute are the reference's.

Also the full-block files of put_block.f (PUT_full_block_OPEN/PUT_full_block,
:47-376: a y-slab with the dissipation diagnostics and the density) and of
put_block2.f (PUT_full_block2_OPEN/PUT_full_block2, :47-362: a sub-box).

The snapshot's kxt, kyt, kzt are MIXING's diffusivities
(mixing80-revised2014.f:22-59) formed on the host from the end-of-step ss and
ss_v, the values MIXING used (nothing after VISCOSITY3D changes them); Ri and
the full blocks' epsh, epsv, chih, chiv come from the context's diagnostics
(hc_set_diagnostics), which the reference forms inside VISCOSITY3D and MIXING
from the step's intermediate state; rrho = rhop + rho_bkg(k) as DENSITY forms
it.  cc1-cc4 are not written: the reference's allocation of those tracers is
commented out (modules80-revised2014.f:245-248), so their toggles (9-12) cannot
be on.  In put_block1.f the u slice (nx+3 wide) is written with a count of
nx+2, which skews its rows; here u keeps its full ux = nx+3 width.
"""
import numpy as np

import hotcross as H
from ncwrite import NcWriter

MISSING = np.float32(-1.0e34)
MESSAGE = "HotCross Version 1.0 - January, 1998, created by J.W.Lavelle, NOAA/PMEL, Seattle"
HISTORY1 = "Non-hydrostatic model of convection occuring in a cross flow"
HISTORY2 = "Ref: Buoyancy Driven Plumes in Rotating, Strat. Cross Flows..., J. G. R., 1996"

# (file name, field, staggering, units, long name): write_snapshota.f:113-170
VARIABLES = (
    ("u", "u", "u", "ms-1", "u-velocity"),
    ("v", "v", "v", "ms-1", "v-velocity"),
    ("w", "w", "w", "ms-1", "vertical velocity"),
    ("p", "p", "s", "newtons/m2", "Pressure Anomaly"),
    ("ss", "ss", "s", "m2s-1", "Horz Mixing Coef."),
    ("ss_v", "ss_v", "s", "m2s-1", "Vert. Shear Mixing"),
    ("temp", "t", "s", "oC", "Potential Temperature"),
    ("s", "s", "s", "o/oo", "Salinity"),
    ("c", "c", "s", "mass_units/m3", "Tracer-0"),
)
DIMS = {"u": ("sz", "sy", "ux"), "v": ("sz", "vy", "sx"), "w": ("wz", "sy", "sx"), "s": ("sz", "sy", "sx")}


def fill_edges(a, lo, box_lo, box_hi):
    """fill_value (hotcross80-revised2014.f:1499-1528): the cells on the
    twelve edges of the box box_lo..box_hi get -1e34; a holds the part of
    that box starting at index lo."""
    on = [np.isin(np.arange(lo[d], lo[d] + a.shape[d]), (box_lo[d], box_hi[d])) for d in range(3)]
    e0, e1, e2 = on[0][:, None, None], on[1][None, :, None], on[2][None, None, :]
    a[(e0 & e1) | (e0 & e2) | (e1 & e2)] = MISSING
    return a


def field_f32(ctx, name, j0=None, j1=None):
    """The field as the writers store it, over j0..j1 (default: all j of its
    bounds), x first."""
    lo, hi = ctx.bounds(name)
    lo, hi = list(lo), list(hi)
    if j0 is not None:
        lo[1] = j0 - 1 if name == "v" else j0
        hi[1] = j1
    if name == "w":
        # ww = -w over (0:nx+1, 0:ny+1, -1:nz+1), edges filled, then sliced
        blo, bhi = ctx.bounds("w")
        return fill_edges(ctx.download_f32("w", -1.0, lo, hi), lo, blo, bhi)
    return ctx.download_f32(name, 1.0, lo, hi)


def _header(ctx, ny_s, ny_v, title, message):
    p = ctx.prm
    dims = [("time", None), ("sx", p.nx + 2), ("sy", ny_s), ("sz", p.nz + 2), ("ux", p.nx + 3),
            ("vy", ny_v), ("wz", p.nz + 3)]
    # character*80 attributes, blank padded (write_snapshota.f:51-55)
    c80 = lambda t: (t or "")[:80].ljust(80)
    gatts = {"title": c80(title), "message": c80(message or MESSAGE), "history0": c80(""),
             "history1": c80(HISTORY1), "history2": c80(HISTORY2)}
    axes = [("time", "f8", ("time",), {"units": "sec"}),
            ("sx", "f4", ("sx",), {"units": "m"}), ("sy", "f4", ("sy",), {"units": "m"}),
            ("sz", "f4", ("sz",), {"units": "m", "positive": "down"}),
            ("ux", "f4", ("ux",), {"units": "m"}), ("vy", "f4", ("vy",), {"units": "m"}),
            ("wz", "f4", ("wz",), {"units": "m", "positive": "down"})]
    return dims, gatts, axes


def _axes(ctx, w, j0=0, jv0=-1):
    """load_axis (netcdf_augmentation.f:62-73): start + d*(i-1), with
    xstart = -nx/3 dx, ystart = -ny/2 dy (hotcross80-revised2014.f:137-138)
    and the staggered starts of write_snapshota.f:247-256."""
    p = ctx.prm
    xstart, ystart = -(p.nx / 3.0) * p.dx, -(p.ny / 2.0) * p.dy
    ax = lambda n, start, d: (d * np.arange(n) + start).astype(np.float32)
    w.put("sx", ax(p.nx + 2, xstart - 0.5 * p.dx, p.dx))
    w.put("sy", ax(w.dimlen["sy"], ystart - 0.5 * p.dy + j0 * p.dy, p.dy))
    w.put("sz", ax(p.nz + 2, p.zstart - 0.5 * p.dz, p.dz))
    w.put("ux", ax(p.nx + 3, xstart - p.dx, p.dx))
    w.put("vy", ax(w.dimlen["vy"], ystart - p.dy + (jv0 + 1) * p.dy, p.dy))
    w.put("wz", ax(p.nz + 3, p.zstart - p.dz, p.dz))


def mixing_diffusivities(ctx, ss, ssv):
    """kxt (-1:nx+1, 0:ny+1, 0:nz+1), kyt (0:nx+1, -1:ny+1, 0:nz+1) and kzt
    (0:nx+1, 0:ny+1, -1:nz+1) of MIXING (mixing80-revised2014.f:22-59) from
    FP64 ss, ss_v over (0:nx+1, 0:ny+1, 0:nz+1), zero where MIXING leaves
    them zero, in its evaluation order."""
    p = ctx.prm
    nx, ny, nz = p.nx, p.ny, p.nz
    pr = p.prandtl
    _, _, _, azzmin = ctx.profiles()
    kxt = np.zeros((nx + 3, ny + 2, nz + 2))
    kyt = np.zeros((nx + 2, ny + 3, nz + 2))
    kzt = np.zeros((nx + 2, ny + 2, nz + 3))
    # kxt(i, j, k), i = 0..nx, j = 0..ny+1, k = 1..nz (array index i+1)
    kxt[1:nx + 2, :, 1:nz + 1] = p.bit5k * ((0.5 * (ss[0:nx + 1, :, 1:nz + 1] + ss[1:nx + 2, :, 1:nz + 1])) / pr
                                           + p.axmin / pr)
    kyt[:, 1:ny + 2, 1:nz + 1] = p.bit5k * ((0.5 * (ss[:, 0:ny + 1, 1:nz + 1] + ss[:, 1:ny + 2, 1:nz + 1])) / pr
                                           + p.aymin / pr)
    az = azzmin[1:nz][None, None, :]
    kzt[:, :, 2:nz + 1] = p.bit2k * ((0.5 * (ssv[:, :, 1:nz] + ssv[:, :, 2:nz + 1])) / pr + az / pr)
    return kxt, kyt, kzt


def sponge_masks(ctx):
    """alphaee (0:nx+1, 0:ny+1, 0:nz+1) = max(alphae, (alphaw(k) +
    alphaw(k-1))/2) with alphae before its periodic copies
    (sponge80.f:107-123), alphau (-1:nx+1, 0:ny+1), alphav (0:nx+1, -1:ny+1)
    and alphae (0:nx+1, 0:ny+1)."""
    au, av, ae, aw = ctx.sponge()
    ae0 = np.maximum(au[1:, :] + au[:-1, :], av[:, 1:] + av[:, :-1]) / 2.0
    lid = (aw[1:] + aw[:-1]) / 2.0  # k = 0..nz+1: (alphaw(k) + alphaw(k-1)) / 2
    aee = np.maximum(ae0[:, :, None], lid[None, None, :])
    return aee, au, av, ae


def _diag_f32(ctx, name, j0, j1):
    if not getattr(ctx, "diagnostics_on", False):
        raise ValueError("%s needs the context's diagnostics (ctx.set_diagnostics(True) before the step)" % name)
    return ctx.diagnostic(name)[:, j0:j1 + 1, :].astype(np.float32)


def _rrho_f32(ctx, j0, j1):
    """rrho = rhop + rho_bkg(k) where DENSITY writes it, the interior and the
    six faces (potential_density80.f:20-67); the box's edges and corners,
    which it never writes, hold zero."""
    rhop = ctx.download("rhop")
    _, _, rho_bkg, _ = ctx.profiles()
    r = rhop + rho_bkg[None, None, :]
    ng = sum(np.isin(np.arange(n), (0, n - 1)).reshape([-1 if d == q else 1 for d in range(3)]).astype(int)
             for q, n in enumerate(r.shape))
    r[ng >= 2] = 0.0
    return r[:, j0:j1 + 1, :].astype(np.float32)


def _toggled(ctx, fields):
    p = ctx.prm
    off = {"temp": p.toggle_t != 1, "s": p.toggle_s != 1, "c": p.toggle_c != 1}
    names = [v[0] for v in VARIABLES if not off.get(v[0], False)]
    return names if fields is None else [n for n in names if n in fields]


def _att(units, lname):
    return {"units": units, "long_name": lname, "missing_value": MISSING}


def snapshot(ctx, path, title="", message=None, fields=None, az_bkg=True):
    """PUT_snapshot_open + PUT_snapshot + PUT_snapshot_close
    (write_snapshota.f): one record of the whole box, the variables in the
    reference's order.  c, kxt, kyt, kzt, alphaee, alphau and alphav are
    fixed variables, as the reference defines them without the time
    dimension (write_snapshota.f:158-240); ri (with ss and ss_v) is written
    when the context's diagnostics are on, Az_bkg (azzmin, toggle 13) when
    az_bkg is set."""
    p = ctx.prm
    names = _toggled(ctx, fields)
    with_ri = "ss" in names and getattr(ctx, "diagnostics_on", False)
    dims, gatts, axes = _header(ctx, p.ny + 2, p.ny + 3, title, message)
    vs = list(axes)
    for fname, field, stag, units, lname in VARIABLES:
        if fname in names:
            d = DIMS[stag] if fname == "c" else ("time",) + DIMS[stag]
            vs.append((fname, "f4", d, _att(units, lname)))
        if fname == "ss_v" and with_ri:
            vs.append(("ri", "f4", ("time",) + DIMS["s"], _att(" ", "Richardson Number")))
    vs += [("kxt", "f4", DIMS["u"], _att("m2/s", "KXT")), ("kyt", "f4", DIMS["v"], _att("m2/s", "KYT")),
           ("kzt", "f4", DIMS["w"], _att("m2/s", "KZT"))]
    if az_bkg:
        vs.append(("Az_bkg", "f4", ("sz",), _att("m2/s", "Background Az")))
    vs += [("alphaee", "f4", ("sz", "sy", "sx"), _att("s^(-1)", "eta_sponge")),
           ("alphau", "f4", ("sy", "ux"), _att("s^(-1)", "u_sponge")),
           ("alphav", "f4", ("vy", "sx"), _att("s^(-1)", "v_sponge"))]
    time, _, _ = ctx.clock()
    with NcWriter(path, dims, vs, gatts) as w:
        _axes(ctx, w)
        rec = {"time": np.array([time])}
        for fname, field, _, _, _ in VARIABLES:
            if fname not in names:
                continue
            a = field_f32(ctx, field)
            if fname == "c":
                w.put("c", a)
            else:
                rec[fname] = a
        if with_ri:
            rec["ri"] = _diag_f32(ctx, "ri", 0, p.ny + 1)
        kxt, kyt, kzt = mixing_diffusivities(ctx, ctx.download("ss"), ctx.download("ss_v"))
        w.put("kxt", kxt.astype(np.float32))
        w.put("kyt", kyt.astype(np.float32))
        w.put("kzt", kzt.astype(np.float32))
        if az_bkg:
            w.put("Az_bkg", ctx.profiles()[3].astype(np.float32))
        aee, au, av, _ = sponge_masks(ctx)
        w.put("alphaee", aee.astype(np.float32))
        w.put("alphau", au.astype(np.float32))
        w.put("alphav", av.astype(np.float32))
        w.append(rec)


class SlabWriter:
    """PUT_full_block1_OPEN / PUT_full_block1 (put_block1.f): one record of
    the y-slab j = nys..nye per call of append(), at the model's time; the
    1 Hz output of the plume runs."""

    def __init__(self, ctx, path, nys, nye, title="", message=None, fields=None):
        p = ctx.prm
        if not (0 <= nys <= nye <= p.ny + 1):
            raise ValueError("slab must satisfy 0 <= nys <= nye <= ny+1")
        self.ctx, self.nys, self.nye = ctx, nys, nye
        self.names = _toggled(ctx, fields)
        dims, gatts, axes = _header(ctx, nye - nys + 1, nye - nys + 2, title, message)
        vs = list(axes)
        for fname, field, stag, units, lname in VARIABLES:
            if fname in self.names:
                vs.append((fname, "f4", ("time",) + DIMS[stag],
                           {"units": units, "long_name": lname, "missing_value": MISSING}))
        self.w = NcWriter(path, dims, vs, gatts)
        _axes(ctx, self.w, nys, nys - 1)

    def append(self):
        time, _, _ = self.ctx.clock()
        rec = {"time": np.array([time])}
        for fname, field, _, _, _ in VARIABLES:
            if fname in self.names:
                rec[fname] = field_f32(self.ctx, field, self.nys, self.nye)
        self.w.append(rec)

    def close(self):
        self.w.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# the dissipation and density variables of the full blocks (put_block.f:226-254)
BLOCK_EXTRAS = (("epsh", "m2/s3", "Horiz. Diss."), ("epsv", "m2/s3", "Vert. Diss."),
                ("chih", "m2/s3", "horz temp var diss"), ("chiv", "m2/s3", "vert temp var diss"),
                ("Ri", " ", "Richardson number"), ("rrho", "kg/m3", "Density"))
_BLOCK_DIAG = {"epsh": "epsh", "epsv": "epsv", "chih": "chi_h", "chiv": "chi_v", "Ri": "ri"}


class FullBlockWriter:
    """PUT_full_block_OPEN / PUT_full_block (put_block.f:47-376): the y-slab
    j = nys..nye over the whole x and z range, one record per append(): the
    SlabWriter's variables plus, with ss and ss_v, the dissipations epsh,
    epsv, chih, chiv, the Richardson number and the density rrho (these
    need the context's diagnostics), and, at open, the sponge masks alphae,
    alphau, alphav of the slab; Az_bkg with the first record (toggle 13)."""

    def __init__(self, ctx, path, nys, nye, title="", message=None, fields=None, az_bkg=True):
        p = ctx.prm
        if not (0 <= nys <= nye <= p.ny + 1):
            raise ValueError("slab must satisfy 0 <= nys <= nye <= ny+1")
        self.ctx, self.nys, self.nye, self.az, self.first = ctx, nys, nye, az_bkg, True
        self.names = _toggled(ctx, fields)
        dims, gatts, axes = _header(ctx, nye - nys + 1, nye - nys + 2, title, message)
        vs = list(axes)
        for fname, field, stag, units, lname in VARIABLES:
            if fname in self.names:
                vs.append((fname, "f4", ("time",) + DIMS[stag], _att(units, lname)))
        if az_bkg:
            vs.append(("Az_bkg", "f4", ("sz",), _att("m2/s", "Background Az")))
        vs += [("alphae", "f4", ("sy", "sx"), _att("s^(-1)", "eta_sponge")),
               ("alphau", "f4", ("sy", "ux"), _att("s^(-1)", "u_sponge")),
               ("alphav", "f4", ("vy", "sx"), _att("s^(-1)", "v_sponge"))]
        self.extras = "ss" in self.names
        if self.extras:
            vs += [(n, "f4", ("time",) + DIMS["s"], _att(u, l)) for n, u, l in BLOCK_EXTRAS]
        self.w = NcWriter(path, dims, vs, gatts)
        _axes(self.ctx, self.w, nys, nys - 1)
        _, au, av, ae = sponge_masks(ctx)
        self.w.put("alphae", ae[:, nys:nye + 1].astype(np.float32))
        self.w.put("alphau", au[:, nys:nye + 1].astype(np.float32))
        self.w.put("alphav", av[:, nys:nye + 2].astype(np.float32))  # (0:nx+1, nys-1:nye)

    def append(self):
        time, _, _ = self.ctx.clock()
        rec = {"time": np.array([time])}
        for fname, field, _, _, _ in VARIABLES:
            if fname in self.names:
                rec[fname] = field_f32(self.ctx, field, self.nys, self.nye)
        if self.extras:
            for n, d in _BLOCK_DIAG.items():
                rec[n] = _diag_f32(self.ctx, d, self.nys, self.nye)
            rec["rrho"] = _rrho_f32(self.ctx, self.nys, self.nye)
        if self.az and self.first:
            self.w.put("Az_bkg", self.ctx.profiles()[3].astype(np.float32))
        self.first = False
        self.w.append(rec)

    def close(self):
        self.w.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class BoxWriter:
    """PUT_full_block2_OPEN / PUT_full_block2 (put_block2.f:47-362): the
    sub-box nxs..nxe x nys..nye x nzs..nze, one record per append(), with
    the variables put_block2.f defines: u, v, w (-w, edges filled), temp,
    s, c, the dissipations, Ri and rrho; as there, c is defined but never
    written (its write is gated on toggle(8) = -1), so it holds netCDF's
    default fill value, and p, ss, ss_v and the sponge masks are absent."""

    FILL = np.float32(9.9692099683868690e36)  # NC_FILL_FLOAT

    def __init__(self, ctx, path, box, title="", message=None):
        p = ctx.prm
        (nxs, nxe), (nys, nye), (nzs, nze) = box
        for lo, hi, n in ((nxs, nxe, p.nx), (nys, nye, p.ny), (nzs, nze, p.nz)):
            if not (0 <= lo <= hi <= n + 1):
                raise ValueError("box must lie in 0..n+1 in every direction")
        self.ctx, self.box = ctx, box
        c80 = lambda t: (t or "")[:80].ljust(80)
        gatts = {"title": c80(title), "message": c80(message or MESSAGE), "history0": c80(""),
                 "history1": c80(HISTORY1), "history2": c80(HISTORY2)}
        dims = [("time", None), ("sx", nxe - nxs + 1), ("sy", nye - nys + 1), ("sz", nze - nzs + 1),
                ("ux", nxe - nxs + 2), ("vy", nye - nys + 2), ("wz", nze - nzs + 2)]
        vs = [("time", "f8", ("time",), {"units": "sec"}),
              ("sx", "f4", ("sx",), {"units": "m"}), ("sy", "f4", ("sy",), {"units": "m"}),
              ("sz", "f4", ("sz",), {"units": "m", "positive": "down"}),
              ("ux", "f4", ("ux",), {"units": "m"}), ("vy", "f4", ("vy",), {"units": "m"}),
              ("wz", "f4", ("wz",), {"units": "m", "positive": "down"}),
              ("u", "f4", ("time",) + DIMS["u"], _att("ms-1", "u-velocity")),
              ("v", "f4", ("time",) + DIMS["v"], _att("ms-1", "v-velocity")),
              ("w", "f4", ("time",) + DIMS["w"], _att("ms-1", "vertical velocity")),
              ("temp", "f4", ("time",) + DIMS["s"], _att("oC", "Potential Temperature")),
              ("s", "f4", ("time",) + DIMS["s"], _att("o/oo", "Salinity")),
              ("c", "f4", ("time",) + DIMS["s"], _att("mass_units/m3", "Tracer-0"))]
        vs += [(n, "f4", ("time",) + DIMS["s"], _att(u, l)) for n, u, l in BLOCK_EXTRAS]
        self.w = NcWriter(path, dims, vs, gatts)
        # load_axis with the box's starts (put_block2.f:250-263)
        xstart, ystart = -(p.nx / 3.0) * p.dx, -(p.ny / 2.0) * p.dy
        ax = lambda n, start, d: (d * np.arange(n) + start).astype(np.float32)
        self.w.put("sx", ax(nxe - nxs + 1, xstart - 0.5 * p.dx + nxs * p.dx, p.dx))
        self.w.put("sy", ax(nye - nys + 1, ystart - 0.5 * p.dy + nys * p.dy, p.dy))
        self.w.put("sz", ax(nze - nzs + 1, p.zstart - 0.5 * p.dz + nzs * p.dz, p.dz))
        self.w.put("ux", ax(nxe - nxs + 2, xstart - p.dx + nxs * p.dx, p.dx))
        self.w.put("vy", ax(nye - nys + 2, ystart - p.dy + nys * p.dy, p.dy))
        self.w.put("wz", ax(nze - nzs + 2, p.zstart - p.dz + nzs * p.dz, p.dz))

    def append(self):
        (nxs, nxe), (nys, nye), (nzs, nze) = self.box
        ctx = self.ctx
        time, _, _ = ctx.clock()
        rec = {"time": np.array([time])}
        rec["u"] = ctx.download_f32("u", 1.0, (nxs - 1, nys, nzs), (nxe, nye, nze))
        rec["v"] = ctx.download_f32("v", 1.0, (nxs, nys - 1, nzs), (nxe, nye, nze))
        blo, bhi = ctx.bounds("w")
        lo = (nxs, nys, nzs - 1)
        rec["w"] = fill_edges(ctx.download_f32("w", -1.0, lo, (nxe, nye, nze)), lo, blo, bhi)
        rec["temp"] = ctx.download_f32("t", 1.0, (nxs, nys, nzs), (nxe, nye, nze))
        rec["s"] = ctx.download_f32("s", 1.0, (nxs, nys, nzs), (nxe, nye, nze))
        rec["c"] = np.full((nxe - nxs + 1, nye - nys + 1, nze - nzs + 1), self.FILL, dtype=np.float32)
        for n, d in _BLOCK_DIAG.items():
            rec[n] = _diag_f32(ctx, d, nys, nye)[nxs:nxe + 1, :, nzs:nze + 1]
        rec["rrho"] = _rrho_f32(ctx, nys, nye)[nxs:nxe + 1, :, nzs:nze + 1]
        self.w.append(rec)

    def close(self):
        self.w.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run(ctx, nsteps, emergency_path=None, chunk=10, on_chunk=None):
    """nsteps of the native loop; on the CFL or NaN bailout (HC_ECFL,
    HC_ENAN) write the emergency snapshot to emergency_path first, as
    CFL_VALUES does (hotcross80-revised2014.f:406-414), then raise.
    on_chunk(ctx) runs after every `chunk` steps (output, statistics)."""
    done = 0
    while done < nsteps:
        n = min(chunk, nsteps - done)
        try:
            ctx.step(n)
        except H.HcError as e:
            if e.code in (H.HC_ECFL, H.HC_ENAN) and emergency_path:
                _, mtime, _ = ctx.clock()
                snapshot(ctx, emergency_path, message="EMERGENCY BAILOUT AT TIME STEP = %d" % mtime)
            raise
        done += n
        if on_chunk:
            on_chunk(ctx)
    return done
