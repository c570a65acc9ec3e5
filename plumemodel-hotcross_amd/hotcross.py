"""Python host harness over the HotCross C-ABI (libhotcross.so, include/hotcross.h).

This mirrors the reference's Fortran interface for the timestep: one method per
reference subroutine (hotcross80-revised2014.f:264-350), module arrays as
Fortran-ordered numpy arrays with the reference's lower bounds, and errors
raised where the reference would `stop`.  The HIP library must be built
(plumemodel-hotcross_amd/build.py); there is no fallback path.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhotcross.so")

FIELDS = ("u", "v", "w", "up", "vp", "wp", "um", "vm", "wm", "p", "t", "s", "c", "ss",
          "ss_v", "rhop", "fx", "fy", "fz")
FIELD_ID = {n: q for q, n in enumerate(FIELDS)}
# adjoint-only fields: the work velocities' Courant numbers (u-, v-, w-bounds)
ADJ_ID = dict(FIELD_ID, uc=32, vc=33, wc=34)
ADJ_BOUNDS = {"uc": "u", "vc": "v", "wc": "w"}
FP64, FP32 = 0, 1
# error codes of include/hotcross.h
HC_EINVAL, HC_EDEVICE, HC_ENOMEM, HC_ECFL, HC_ENAN = -1, -2, -3, -4, -5


class HcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class HcParams(ctypes.Structure):
    """hc_params of include/hotcross.h."""
    _fields_ = (
        [(n, ctypes.c_int32) for n in ("nx", "ny", "nz", "nt")]
        + [(n, ctypes.c_double) for n in ("dx", "dy", "dz", "dt", "zstart", "zend")]
        + [(n, ctypes.c_int32) for n in (
            "bit2", "bit3", "bit4", "bit5", "bit7", "bit8", "bit9", "bit10",
            "bit2k", "bit5k", "ismol", "ismag_h", "ismag_v", "bcondx", "bcondy",
            "bottom_slip", "side_slip", "open_box")]
        + [("bdbit", ctypes.c_int32 * 4)]
        + [(n, ctypes.c_int32) for n in ("modulo_x", "modulo_y", "toggle_t",
                                         "toggle_s", "toggle_c", "pad0")]
        + [(n, ctypes.c_double) for n in (
            "cs_h", "cs_v", "prandtl", "axmin", "aymin", "azmin", "azmax",
            "mixing_height_scale", "sponge_factor", "lat", "g", "rho0", "assel",
            "eps", "omega", "source_length", "source_width", "dqtdt", "dqsdt",
            "dqcdt", "tsource", "ssource", "wsource", "source_ramp_hrs",
            "noise_amp", "wramp_hrs", "flow_u0", "flow_u1", "flow_v1",
            "flow_omega", "flow_ramp_hrs")]
        + [(n, ctypes.c_int32) for n in ("eos", "pad1")]
        + [(n, ctypes.c_double) for n in ("eos_rho_ref", "eos_alpha", "eos_beta",
                                          "eos_t0", "eos_s0")]
    )


def field_bounds(name, nx, ny, nz):
    lo = [0, 0, 0]
    hi = [nx + 1, ny + 1, nz + 1]
    if name in ("u", "up", "um", "fx"):
        lo[0] = -1
    elif name in ("v", "vp", "vm", "fy"):
        lo[1] = -1
    elif name in ("w", "wp", "wm", "fz"):
        lo[2] = -1
    return tuple(lo), tuple(hi)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhotcross.so is not built: run plumemodel-hotcross_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, dbl, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p
        sig = {
            "hc_create": [ctypes.POINTER(HcParams), ctypes.c_int, ctypes.POINTER(vp)],
            "hc_destroy": [vp], "hc_synchronize": [vp], "hc_stream": [vp],
            "hc_upload": [vp, ctypes.c_int, dp, dp, dp],
            "hc_download": [vp, ctypes.c_int, dp, dp, dp],
            "hc_set_profiles": [vp, dp, dp, dp, dp],
            "hc_set_sponge": [vp, dp, dp, dp, dp],
            "hc_get_profiles": [vp, dp, dp, dp, dp],
            "hc_get_sponge": [vp, dp, dp, dp, dp],
            "hc_set_diagnostics": [vp, i32],
            "hc_get_diagnostic": [vp, i32, dp],
            "hc_init_case": [vp, dp, dp, dp, i32, dbl, ctypes.c_uint64],
            "hc_begin_step": [vp, i32, dbl],
            "hc_body_force": [vp], "hc_forces": [vp], "hc_pressure": [vp],
            "hc_momentum_u": [vp], "hc_momentum_v": [vp], "hc_momentum_w": [vp],
            "hc_viscosity3d": [vp], "hc_source_terms": [vp, dp], "hc_mixing": [vp],
            "hc_work_velocity": [vp, dbl],
            "hc_tracer31": [vp, ctypes.c_int, dp, dbl, dbl, dbl, dp],
            "hc_vent_ghost": [vp, ctypes.c_int, dbl],
            "hc_density": [vp], "hc_refresh": [vp], "hc_cfl_values": [vp, dp],
            "hc_step": [vp, i32], "hc_source_noise": [vp, i32, dp],
            "hc_stats_reset": [vp, i32, i32, i32], "hc_stats_accumulate": [vp],
            "hc_stats_get": [vp, dp, dp],
            "hc_clock": [vp, dp, dp, dp],
            "hc_download_f32": [vp, ctypes.c_int, dp, dp, dp, dbl],
            "hc_phase_timing": [vp, ctypes.c_int],
            "hc_phase_times": [vp, ctypes.POINTER(ctypes.c_char_p), dp, dp, i32],
            "hc_set_controls": [vp, dp, dbl, dbl],
            "hc_adj_reset": [vp],
            "hc_adj_upload": [vp, ctypes.c_int, dp, dp, dp],
            "hc_adj_download": [vp, ctypes.c_int, dp, dp, dp],
            "hc_adj_forces": [vp], "hc_adj_pressure": [vp],
            "hc_adj_momentum_u": [vp], "hc_adj_momentum_v": [vp], "hc_adj_momentum_w": [vp],
            "hc_adj_viscosity3d": [vp], "hc_adj_work_velocity": [vp, dbl],
            "hc_adj_tracer31": [vp, ctypes.c_int, dp, dbl, dbl, dbl],
            "hc_adj_vent_ghost": [vp, ctypes.c_int],
            "hc_adj_density": [vp], "hc_adj_refresh": [vp],
            "hc_adj_gradients": [vp, dp, dp],
            "hc_adj_set_probes": [vp, i32, dp, i32, dbl, dbl],
            "hc_adj_window": [vp, i32, i32, dp, dp, dp, dp, dp, i32],
            "hc_adj_window_bytes": [vp, i32, i32, dp],
            "hc_adj_window_binomial": [vp, i32, i32, dp, dp, dp, dp, dp, i32],
            "hc_adj_window_binomial_bytes": [vp, i32, i32, dp],
            "hc_binomial_schedule": [i32, i32, vp, ctypes.c_int64, vp, vp],
        }
        for n, a in sig.items():
            f = getattr(L, n)
            f.argtypes = a
            f.restype = ctypes.c_int
        L.hc_stream.restype = ctypes.c_void_p
        L.hc_last_error.argtypes = [vp]
        L.hc_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


# hc_get_diagnostic's order
DIAGNOSTICS = ("ri", "epsh", "epsv", "chi_v", "chi_h")


def _arr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# sums of hc_stats_get, in order (include/hotcross.h)
NSTAT = 10
STAT_NAMES = ("w", "t", "uc", "vc", "wc", "uc2", "vc2", "wc2", "t2", "wct")


def plane_statistics(sums, count):
    """Time means from the running sums of a statistics slab.  Returns a dict
    of arrays shaped (nk, ny, nx): w_up (the time-mean vertical velocity at
    the w points, positive upward as the writers output it,
    put_block1.f:265), t (mean temperature), tke (0.5 times the summed
    variances of the cell-centred velocity components), t_var (temperature
    variance) and wt_flux (<w'T'> with w positive upward)."""
    if count < 1:
        raise ValueError("no samples accumulated")
    m = {n: sums[q] / count for q, n in enumerate(STAT_NAMES)}
    var = lambda sq, mean: np.maximum(sq - mean * mean, 0.0)
    return {
        "w_up": -m["w"],
        "t": m["t"],
        "tke": 0.5 * (var(m["uc2"], m["uc"]) + var(m["vc2"], m["vc"]) + var(m["wc2"], m["wc"])),
        "t_var": var(m["t2"], m["t"]),
        "wt_flux": -(m["wct"] - m["wc"] * m["t"]),
    }


class Context:
    """One device state of the timestep (a context of the C-ABI)."""

    def __init__(self, prm, precision=FP64):
        self.L = lib()
        self.prm = prm
        self.nx, self.ny, self.nz = prm.nx, prm.ny, prm.nz
        self.h = ctypes.c_void_p()
        r = self.L.hc_create(ctypes.byref(prm), precision, ctypes.byref(self.h))
        self._check(r)
        self.time = -prm.dt
        self.m = 0
        self.diagnostics_on = False

    def _check(self, r):
        if r != 0:
            msg = self.L.hc_last_error(self.h).decode() if self.h else "hc_create failed"
            raise HcError(r, msg)
        return r

    def close(self):
        if self.h:
            self.L.hc_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state transfer ----------------------------------------------------
    def bounds(self, name):
        return field_bounds(name, self.nx, self.ny, self.nz)

    def upload(self, name, arr):
        lo, hi = self.bounds(name)
        shape = tuple(h - l + 1 for l, h in zip(lo, hi))
        a = np.asfortranarray(np.asarray(arr, dtype=np.float64).reshape(shape, order="F"))
        lo_a = np.array(lo, dtype=np.int32)
        hi_a = np.array(hi, dtype=np.int32)
        self._check(self.L.hc_upload(self.h, FIELD_ID[name], _arr(a), _arr(lo_a), _arr(hi_a)))

    def download(self, name):
        lo, hi = self.bounds(name)
        shape = tuple(h - l + 1 for l, h in zip(lo, hi))
        a = np.zeros(shape, dtype=np.float64, order="F")
        lo_a = np.array(lo, dtype=np.int32)
        hi_a = np.array(hi, dtype=np.int32)
        self._check(self.L.hc_download(self.h, FIELD_ID[name], _arr(a), _arr(lo_a), _arr(hi_a)))
        return a

    def download_f32(self, name, scale=1.0, lo=None, hi=None):
        """FP32 copy of scale * field over lo..hi (default: the field's
        bounds), converted on the device; Fortran-ordered like download."""
        blo, bhi = self.bounds(name)
        lo = blo if lo is None else lo
        hi = bhi if hi is None else hi
        shape = tuple(h - l + 1 for l, h in zip(lo, hi))
        a = np.zeros(shape, dtype=np.float32, order="F")
        lo_a = np.array(lo, dtype=np.int32)
        hi_a = np.array(hi, dtype=np.int32)
        self._check(self.L.hc_download_f32(self.h, FIELD_ID[name], _arr(a), _arr(lo_a), _arr(hi_a), scale))
        return a

    def clock(self):
        """(time, mtime) of the last step begun and the next step of hc_step."""
        t = np.zeros(1)
        m = np.zeros(2, dtype=np.int32)
        self._check(self.L.hc_clock(self.h, _arr(t), _arr(m[0:1]), _arr(m[1:2])))
        return float(t[0]), int(m[0]), int(m[1])

    def set_profiles(self, t_bkg, s_bkg, rho_bkg, azzmin):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (t_bkg, s_bkg, rho_bkg, azzmin)]
        self._t_bkg, self._s_bkg = arrs[0].copy(), arrs[1].copy()
        self._check(self.L.hc_set_profiles(self.h, *[_arr(x) for x in arrs]))

    def profiles(self):
        """(t_bkg, s_bkg, rho_bkg, azzmin), each over k = 0..nz+1."""
        a = [np.zeros(self.prm.nz + 2) for _ in range(4)]
        self._check(self.L.hc_get_profiles(self.h, *[_arr(x) for x in a]))
        return tuple(a)

    def sponge(self):
        """(alphau, alphav, alphae, alphaw): alphau over (-1:nx+1, 0:ny+1),
        alphav over (0:nx+1, -1:ny+1), alphae over (0:nx+1, 0:ny+1) (x
        first), alphaw over -1..nz+1."""
        p = self.prm
        a = [np.zeros((p.nx + 3) * (p.ny + 2)), np.zeros((p.nx + 2) * (p.ny + 3)), np.zeros((p.nx + 2) * (p.ny + 2)),
             np.zeros(p.nz + 3)]
        self._check(self.L.hc_get_sponge(self.h, *[_arr(x) for x in a]))
        return (a[0].reshape((p.nx + 3, p.ny + 2), order="F"), a[1].reshape((p.nx + 2, p.ny + 3), order="F"),
                a[2].reshape((p.nx + 2, p.ny + 2), order="F"), a[3])

    def set_diagnostics(self, on=True):
        """Form Ri, epsh, epsv, chi_v and chi_h at every later VISCOSITY3D."""
        self._check(self.L.hc_set_diagnostics(self.h, 1 if on else 0))
        self.diagnostics_on = bool(on)

    def diagnostic(self, name):
        """Diagnostic `name` (ri, epsh, epsv, chi_v, chi_h) of the last step,
        FP64 over (0:nx+1, 0:ny+1, 0:nz+1), x first."""
        p = self.prm
        which = DIAGNOSTICS.index(name)
        a = np.zeros((p.nx + 2) * (p.ny + 2) * (p.nz + 2))
        self._check(self.L.hc_get_diagnostic(self.h, which, _arr(a)))
        return a.reshape((p.nx + 2, p.ny + 2, p.nz + 2), order="F")

    def set_sponge(self, alphau, alphav, alphae, alphaw):
        arrs = [np.asfortranarray(x, dtype=np.float64) for x in (alphau, alphav, alphae, alphaw)]
        self._check(self.L.hc_set_sponge(self.h, *[_arr(x) for x in arrs]))

    def init_case(self, z, t, s, t_noise=0.0, seed=1):
        z, t, s = (np.ascontiguousarray(x, dtype=np.float64) for x in (z, t, s))
        self._check(self.L.hc_init_case(self.h, _arr(z), _arr(t), _arr(s), len(z), t_noise, seed))

    # ---- the reference's subroutines --------------------------------------
    def begin_step(self, m, time):
        self._check(self.L.hc_begin_step(self.h, m, time))

    def body_force(self):
        self._check(self.L.hc_body_force(self.h))

    def forces(self):
        self._check(self.L.hc_forces(self.h))

    def pressure(self):
        self._check(self.L.hc_pressure(self.h))

    def momentum_u(self):
        self._check(self.L.hc_momentum_u(self.h))

    def momentum_v(self):
        self._check(self.L.hc_momentum_v(self.h))

    def momentum_w(self):
        self._check(self.L.hc_momentum_w(self.h))

    def viscosity3d(self):
        self._check(self.L.hc_viscosity3d(self.h))

    def source_terms(self, noise=None):
        if noise is None:
            self._check(self.L.hc_source_terms(self.h, None))
        else:
            n = np.ascontiguousarray(noise, dtype=np.float64)
            self._check(self.L.hc_source_terms(self.h, _arr(n)))

    def mixing(self):
        self._check(self.L.hc_mixing(self.h))

    def work_velocity(self, tau3):
        self._check(self.L.hc_work_velocity(self.h, tau3))

    def tracer31(self, name, c_bkg, cmin, cmax, tau3):
        b = np.ascontiguousarray(c_bkg, dtype=np.float64)
        out = np.zeros(3)
        self._check(self.L.hc_tracer31(self.h, FIELD_ID[name], _arr(b), cmin, cmax, tau3, _arr(out)))
        return out

    def vent_ghost(self, name, value):
        self._check(self.L.hc_vent_ghost(self.h, FIELD_ID[name], value))

    def density(self):
        self._check(self.L.hc_density(self.h))

    def refresh(self):
        self._check(self.L.hc_refresh(self.h))

    def cfl_values(self):
        out = np.zeros(6)
        r = self.L.hc_cfl_values(self.h, _arr(out))
        self._check(r)
        return out

    def driver_step(self, m, noise=None):
        """The time loop body (hotcross80-revised2014.f:264-350) through the
        per-subroutine entries, in the reference's order."""
        p = self.prm
        self.time = self.time + p.dt
        self.begin_step(m, self.time)
        self.body_force()
        self.forces()
        self.pressure()
        self.momentum_u()
        self.momentum_v()
        self.momentum_w()
        self.viscosity3d()
        self.source_terms(noise)
        self.mixing()
        if p.toggle_t == 1:
            self.work_velocity(p.dt)
            self.tracer31("t", self._t_bkg, self._t_bkg.min(), 1.0e10, p.dt)
            self.vent_ghost("t", p.tsource + self._t_bkg[p.nz])
        if p.toggle_s == 1:
            self.tracer31("s", self._s_bkg, 0.0, self._s_bkg.max(), p.dt)
            self.vent_ghost("s", p.ssource + self._s_bkg[p.nz])
        if p.bit3 == 1:
            self.density()
        if p.toggle_c == 1:
            self.tracer31("c", np.zeros(p.nz + 2), 0.0, 1.0e10, p.dt)
        self.refresh()
        if m % 10 == 0:
            return self.cfl_values()
        return None

    def step(self, n):
        """n steps of the native loop (hc_step)."""
        self._check(self.L.hc_step(self.h, n))

    def synchronize(self):
        self._check(self.L.hc_synchronize(self.h))

    def stream(self):
        return self.L.hc_stream(self.h)

    def source_noise(self, m):
        out = np.zeros(3 * 4096)
        n = self.L.hc_source_noise(self.h, m, _arr(out))
        return out[: 3 * n].copy()

    def npatch(self):
        return len(self.source_noise(0)) // 3

    # ---- plane statistics ------------------------------------------------------
    def stats_reset(self, k0, k1=None, every=0):
        """Zero the running sums over planes k0..k1 (k1 defaults to k0); the
        native loop adds a sample after every `every`-th step (0: only
        stats_accumulate adds)."""
        k1 = k0 if k1 is None else k1
        self._check(self.L.hc_stats_reset(self.h, k0, k1, every))
        self._stats_k = (k0, k1)

    def stats_accumulate(self):
        self._check(self.L.hc_stats_accumulate(self.h))

    def stats_get(self):
        """(sums with shape (NSTAT, k1-k0+1, ny, nx), sample count)."""
        k0, k1 = self._stats_k
        sums = np.zeros((NSTAT, k1 - k0 + 1, self.ny, self.nx))
        cnt = np.zeros(1, dtype=np.int32)
        self._check(self.L.hc_stats_get(self.h, _arr(sums), _arr(cnt)))
        return sums, int(cnt[0])

    def stats(self):
        """Time means of the slab (plane_statistics of stats_get)."""
        return plane_statistics(*self.stats_get())

    # ---- controls and discrete adjoint ---------------------------------------
    def set_controls(self, qmap=None, flow_scale=1.0, cs=0.0):
        """Heat-flux multiplier per source cell (None = 1), crossflow
        amplitude factor, Smagorinsky constant (cs <= 0 keeps the case's)."""
        q = None if qmap is None else np.ascontiguousarray(qmap, dtype=np.float64)
        self._check(self.L.hc_set_controls(self.h, None if q is None else _arr(q), float(flow_scale), float(cs)))

    def adj_reset(self):
        self._check(self.L.hc_adj_reset(self.h))

    def _adj_box(self, name):
        lo, hi = self.bounds(ADJ_BOUNDS.get(name, name))
        return lo, hi, tuple(h - l + 1 for l, h in zip(lo, hi))

    def adj_upload(self, name, arr):
        lo, hi, shape = self._adj_box(name)
        a = np.asfortranarray(np.asarray(arr, dtype=np.float64).reshape(shape, order="F"))
        self._check(self.L.hc_adj_upload(self.h, ADJ_ID[name], _arr(a), _arr(np.array(lo, dtype=np.int32)),
                                         _arr(np.array(hi, dtype=np.int32))))

    def adj_download(self, name):
        lo, hi, shape = self._adj_box(name)
        a = np.zeros(shape, dtype=np.float64, order="F")
        self._check(self.L.hc_adj_download(self.h, ADJ_ID[name], _arr(a), _arr(np.array(lo, dtype=np.int32)),
                                           _arr(np.array(hi, dtype=np.int32))))
        return a

    def adj_forces(self):
        self._check(self.L.hc_adj_forces(self.h))

    def adj_pressure(self):
        self._check(self.L.hc_adj_pressure(self.h))

    def adj_momentum_u(self):
        self._check(self.L.hc_adj_momentum_u(self.h))

    def adj_momentum_v(self):
        self._check(self.L.hc_adj_momentum_v(self.h))

    def adj_momentum_w(self):
        self._check(self.L.hc_adj_momentum_w(self.h))

    def adj_viscosity3d(self):
        self._check(self.L.hc_adj_viscosity3d(self.h))

    def adj_work_velocity(self, tau3):
        self._check(self.L.hc_adj_work_velocity(self.h, tau3))

    def adj_tracer31(self, name, c_bkg, cmin, cmax, tau3):
        cb = np.ascontiguousarray(c_bkg, dtype=np.float64)
        self._check(self.L.hc_adj_tracer31(self.h, FIELD_ID[name], _arr(cb), cmin, cmax, tau3))

    def adj_vent_ghost(self, name):
        self._check(self.L.hc_adj_vent_ghost(self.h, FIELD_ID[name]))

    def adj_density(self):
        self._check(self.L.hc_adj_density(self.h))

    def adj_refresh(self):
        self._check(self.L.hc_adj_refresh(self.h))

    def adj_gradients(self):
        """(dL/dqmap per source cell, [dL/dflow_scale, dL/dcs_h, dL/dcs_v])."""
        gq = np.zeros(self.npatch())
        out = np.zeros(3)
        self._check(self.L.hc_adj_gradients(self.h, _arr(gq), _arr(out)))
        return gq, out

    def set_probes(self, ijk, every, wgt_w=1.0, wgt_t=1.0):
        a = np.ascontiguousarray(np.asarray(ijk, dtype=np.int32).reshape(-1, 3))
        self._probes = len(a)
        self._every = int(every)
        self._check(self.L.hc_adj_set_probes(self.h, len(a), _arr(a), int(every), float(wgt_w), float(wgt_t)))

    def adj_window_bytes(self, nsteps, seg):
        """HBM bytes of the window's checkpoints and tapes."""
        b = np.zeros(1, dtype=np.uint64)
        self._check(self.L.hc_adj_window_bytes(self.h, nsteps, seg, _arr(b)))
        return int(b[0])

    def adj_window(self, nsteps, seg, obs, end_state=True):
        """Forward nsteps with a checkpoint every seg steps, probe-misfit loss
        against obs (shape (nsteps // every, nprobe, 2): w, T), reverse sweep.
        Returns (loss, samples, dL/dqmap, [dL/dflow_scale, dL/dcs_h, dL/dcs_v]);
        the context then holds the window's end state, or (end_state=False)
        its initial state and step counter."""
        ns = nsteps // self._every
        o = np.ascontiguousarray(np.asarray(obs, dtype=np.float64).reshape(ns * self._probes * 2))
        samp = np.zeros_like(o)
        loss = np.zeros(1)
        gq = np.zeros(self.npatch())
        gs = np.zeros(3)
        self._check(self.L.hc_adj_window(self.h, nsteps, seg, _arr(o), _arr(samp), _arr(loss), _arr(gq), _arr(gs),
                                         1 if end_state else 0))
        return float(loss[0]), samp.reshape(ns, self._probes, 2), gq, gs

    def adj_window_binomial_bytes(self, nsteps, nsnap):
        """HBM bytes of the binomial window's snapshots and tapes."""
        b = np.zeros(1, dtype=np.uint64)
        self._check(self.L.hc_adj_window_binomial_bytes(self.h, nsteps, nsnap, _arr(b)))
        return int(b[0])

    def adj_window_binomial(self, nsteps, nsnap, obs, end_state=True):
        """adj_window under binomial checkpointing with nsnap snapshot slots."""
        ns = nsteps // self._every
        o = np.ascontiguousarray(np.asarray(obs, dtype=np.float64).reshape(ns * self._probes * 2))
        samp = np.zeros_like(o)
        loss = np.zeros(1)
        gq = np.zeros(self.npatch())
        gs = np.zeros(3)
        self._check(self.L.hc_adj_window_binomial(self.h, nsteps, nsnap, _arr(o), _arr(samp), _arr(loss), _arr(gq),
                                                  _arr(gs), 1 if end_state else 0))
        return float(loss[0]), samp.reshape(ns, self._probes, 2), gq, gs

    def phase_timing(self, enable=True):
        self._check(self.L.hc_phase_timing(self.h, 1 if enable else 0))

    def phase_times(self):
        cap = 32
        names = (ctypes.c_char_p * cap)()
        ms = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int32)
        n = self.L.hc_phase_times(self.h, names, _arr(ms), _arr(cnt), cap)
        return {names[q].decode(): (float(ms[q]), int(cnt[q])) for q in range(min(n, cap))}


def binomial_schedule(nsteps, nsnap):
    """The binomial window's schedule, computed on the host (no device):
    (actions as rows (kind, step, slot) with kind 0 advance, 1 snapshot,
    2 restore, 3 turn; the forward steps it runs)."""
    L = lib()
    na, nf = np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64)
    if L.hc_binomial_schedule(nsteps, nsnap, None, 0, _arr(na), _arr(nf)) != 0:
        raise ValueError("nsteps and nsnap must be positive")
    acts = np.zeros((int(na[0]), 4), dtype=np.int32)
    L.hc_binomial_schedule(nsteps, nsnap, _arr(acts), int(na[0]), _arr(na), _arr(nf))
    return acts[:, :3], int(nf[0])


# --------------------------------------------------------------------------
# workloads of the headline measurement
PI = 4.0 * math.atan2(1.0, 1.0)


def workload_params(name="les256"):
    """Case parameters of the named workload.

    les64:   64^3, dx=dy=dz=1 m, dt=0.25 s, uniform 0.05 m/s crossflow.
    les256:  256^3, dx=dy=dz=0.5 m, dt=0.2 s, tidal crossflow
             U(t)=0.03+0.05 sin(2 pi t/44712 s), 50 MW over 4 x 4.5 m with
             +-5% per-cell heat-flux noise, passive tracer.
    les384:  384 x 384 x 320, dx=0.5 m, same physics as les256.
    Shared physics: linear T(z) with N = 1e-3 1/s, S = 34.6, linear equation
    of state, Smagorinsky Cs = 0.2, closed-box y ends with sponge.
    """
    grids = {"les64": (64, 64, 64, 1.0, 0.25), "les256": (256, 256, 256, 0.5, 0.2),
             "les384": (384, 384, 320, 0.5, 0.2), "les128": (128, 128, 128, 1.0, 0.25)}
    nx, ny, nz, d, dt = grids[name]
    p = HcParams()
    p.nx, p.ny, p.nz, p.nt = nx, ny, nz, 3600
    p.dx = p.dy = p.dz = d
    p.dt = dt
    p.zstart = 1750.0
    p.zend = p.zstart + nz * d
    for n in ("bit2", "bit3", "bit4", "bit5", "bit10", "bit2k", "bit5k", "ismol", "ismag_h",
              "ismag_v", "bcondx", "bcondy", "bottom_slip", "side_slip", "open_box",
              "toggle_t", "toggle_s", "toggle_c"):
        setattr(p, n, 1)
    p.bit7, p.bit8, p.bit9 = 0, 0, 0
    p.bdbit[:] = [1, 1, 1, 1]
    p.cs_h = p.cs_v = 0.2
    p.prandtl = 1.0
    p.axmin = p.aymin = 1.0e-4
    p.azmin, p.azmax, p.mixing_height_scale = 1.0e-5, 1.0e-4, 5.0
    p.sponge_factor = 0.05
    p.lat = 47.95
    p.g, p.rho0, p.assel, p.eps = 9.81, 1027.65, 0.01, 1.0e-10
    p.omega = 2.0 * PI / 24.0 / 3600.0
    p.source_length, p.source_width = 4.0, 4.5
    # 50 MW: dQtdt per source cell in K m^3/s (q/(rho0 Cp)/ncells)
    nnx = int(math.floor(4.0 / d + 0.5))
    nny = int(math.floor(4.5 / d + 0.5))
    p.dqtdt = 50.0e6 / (1027.65 * 4200.0) / (nnx * nny)
    p.dqsdt, p.dqcdt = 0.0, 1.0
    p.tsource, p.ssource, p.wsource = 0.0, 0.0, 0.0
    p.source_ramp_hrs, p.noise_amp, p.wramp_hrs = 1.0e-9, 0.05, 6.0
    if name == "les64":
        p.flow_u0, p.flow_u1, p.flow_v1 = 0.05, 0.0, 0.0
        p.noise_amp = 0.0
    else:
        p.flow_u0, p.flow_u1, p.flow_v1 = 0.03, 0.05, 0.0
    p.flow_omega = 2.0 * PI / 44712.0
    p.flow_ramp_hrs = 0.0
    p.eos = 1
    N2 = 1.0e-6
    p.eos_rho_ref, p.eos_alpha, p.eos_beta, p.eos_t0, p.eos_s0 = 1027.65, 2.0e-4, 7.6e-4, 2.0, 34.6
    return p


def linear_stratification(p, n2=1.0e-6, t0=2.0, s0=34.6):
    """(z, T, S) table for INITIALIZE_PROFILES: linear T(z) with buoyancy
    frequency sqrt(n2) under the linear equation of state, constant S; z is
    depth, increasing downward as in the reference's tables."""
    dtdz = n2 / (p.g * p.eos_alpha)  # K per m, warmer upward
    z = np.arange(p.zstart - 4.0, p.zend + 4.0 + 1e-9, 1.0)
    t = t0 + dtdz * (p.zend - z)
    s = np.full_like(z, s0)
    return z, t, s
