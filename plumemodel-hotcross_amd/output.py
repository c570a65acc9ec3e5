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

Not written: ri, kxt/kyt/kzt, cc1-cc4, Az_bkg and the sponge masks, which
the library does not keep as fields (the diffusivities are formed inside the
tracer kernels).  In put_block1.f the u slice (nx+3 wide) is written with a
count of nx+2, which skews its rows; here u keeps its full ux = nx+3 width.
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


def _toggled(ctx, fields):
    p = ctx.prm
    off = {"temp": p.toggle_t != 1, "s": p.toggle_s != 1, "c": p.toggle_c != 1}
    names = [v[0] for v in VARIABLES if not off.get(v[0], False)]
    return names if fields is None else [n for n in names if n in fields]


def snapshot(ctx, path, title="", message=None, fields=None):
    """PUT_snapshot_open + PUT_snapshot + PUT_snapshot_close
    (write_snapshota.f): one record of the whole box.  c is a fixed
    variable over (sz, sy, sx), as the reference defines it with three
    dimensions (write_snapshota.f:158)."""
    p = ctx.prm
    names = _toggled(ctx, fields)
    dims, gatts, axes = _header(ctx, p.ny + 2, p.ny + 3, title, message)
    vs = list(axes)
    for fname, field, stag, units, lname in VARIABLES:
        if fname in names:
            d = DIMS[stag] if fname == "c" else ("time",) + DIMS[stag]
            vs.append((fname, "f4", d, {"units": units, "long_name": lname, "missing_value": MISSING}))
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
