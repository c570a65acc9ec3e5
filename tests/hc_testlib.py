"""Shared helpers for the HotCross tests: case parameters, the oracle, the
compiled reference (oracle/_ref) and array bookkeeping.

The oracle (oracle/liboracle.so) and the reference build are test
infrastructure: they are the checker for the HIP library, never part of it.
"""
import ctypes
import math
import multiprocessing as mp
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
ORACLE_CS_SO = os.path.join(ROOT, "oracle", "liboracle_cs.so")
REF_DIR = "/root/reference"


class Params(ctypes.Structure):
    """Mirror of hc_params in include/hotcross.h."""
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

    def copy(self):
        q = Params()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(self), ctypes.sizeof(Params))
        return q


PI = 4.0 * math.atan2(1.0, 1.0)


def reference_params(nx, ny, nz, dx=1.0, dy=1.0, dz=1.0, dt=0.25, lat=47.95,
                     bit7=1, bcondy=1, zstart=1750.0):
    """A case that uses the reference's hard-coded constants (initialize_
    constants, site_forcing.inc, source80.inc ramps and noise, UNESCO EOS) and
    the switches of a Dante-like run, so the oracle can be compared with the
    compiled reference."""
    p = Params()
    p.nx, p.ny, p.nz, p.nt = nx, ny, nz, 1000
    p.dx, p.dy, p.dz, p.dt = dx, dy, dz, dt
    p.zstart, p.zend = zstart, zstart + nz * dz
    for n in ("bit2", "bit3", "bit4", "bit5", "bit10", "bit2k", "bit5k", "ismol",
              "ismag_h", "ismag_v", "bcondx", "bottom_slip", "side_slip",
              "open_box", "toggle_t", "toggle_s", "toggle_c"):
        setattr(p, n, 1)
    p.bit7, p.bit8, p.bit9 = bit7, 0, 0
    p.bcondy = bcondy
    p.bdbit[:] = [1, 1, 1, 1]
    p.modulo_x = p.modulo_y = 0
    p.cs_h = p.cs_v = 0.2
    p.prandtl = 1.0
    p.axmin = p.aymin = 1.0e-4
    p.azmin, p.azmax, p.mixing_height_scale = 1.0e-5, 1.0e-3, 3.0
    p.sponge_factor = 0.05
    p.lat = lat
    # constants_module (modules80-revised2014.f:51-60)
    p.g, p.rho0, p.assel, p.eps = 9.81, 1027.65, 0.01, 1.0e-10
    p.omega = 2.0 * PI / 24.0 / 3600.0
    p.source_length, p.source_width = 4.0 * dx, 4.5 * dy
    p.dqtdt = 50.0e6 / (1027.65 * 4200.0) / 20.0
    p.dqsdt, p.dqcdt = -1.0e-3, 1.0
    p.tsource, p.ssource, p.wsource = 0.5, -0.01, 0.01
    p.source_ramp_hrs, p.noise_amp, p.wramp_hrs = 2.0, 0.025, 6.0
    p.flow_u0, p.flow_u1, p.flow_v1 = 0.032, 0.036, 0.0024
    p.flow_omega = 2.0 * PI / 12.42 / 3600.0
    p.flow_ramp_hrs = 2.0
    p.eos = 0
    p.eos_rho_ref, p.eos_alpha, p.eos_beta, p.eos_t0, p.eos_s0 = 1027.65, 2.0e-4, 7.6e-4, 2.0, 34.6
    return p


def ref_int_real(p):
    ip = [p.bit2, p.bit3, p.bit4, p.bit5, p.bit7, p.bit8, p.bit9, p.bit10,
          p.bit2k, p.bit5k, p.ismol, p.ismag_h, p.ismag_v, p.bcondx, p.bcondy,
          p.bottom_slip, p.side_slip, p.open_box] + list(p.bdbit) + [
          p.modulo_x, p.modulo_y, p.toggle_t, p.toggle_s, p.toggle_c, p.nt]
    rp = [p.dx, p.dy, p.dz, p.dt, p.zstart, p.zend, p.cs_h, p.cs_v, p.prandtl,
          p.axmin, p.aymin, p.azmin, p.azmax, p.mixing_height_scale,
          p.sponge_factor, p.lat, p.source_length, p.source_width, p.dqtdt,
          p.dqsdt, p.dqcdt, p.tsource, p.ssource, p.wsource]
    return (ctypes.c_int * 28)(*ip), (ctypes.c_double * 24)(*rp)


# --------------------------------------------------------------------------
# array classes (Fortran bounds)
def bounds(name, nx, ny, nz):
    U = ((-1, nx + 1), (0, ny + 1), (0, nz + 1))
    V = ((0, nx + 1), (-1, ny + 1), (0, nz + 1))
    W = ((0, nx + 1), (0, ny + 1), (-1, nz + 1))
    S = ((0, nx + 1), (0, ny + 1), (0, nz + 1))
    table = {
        "u": U, "up": U, "um": U, "fx": U, "kxt": U, "uc": U, "ud": U, "worku": U,
        "v": V, "vp": V, "vm": V, "fy": V, "kyt": V, "vc": V, "vd": V, "workv": V,
        "w": W, "wp": W, "wm": W, "fz": W, "kzt": W, "wc": W, "wd": W, "workw": W,
        "azu": ((-1, nx + 1), (0, ny + 1), (-1, nz + 1)),
        "azv": ((0, nx + 1), (0, ny), (-1, nz + 1)),
        "ayu": ((0, nx), (0, ny), (0, nz + 1)),
        "axv": ((0, nx), (0, ny), (0, nz + 1)),
        "alphau": ((-1, nx + 1), (0, ny + 1)),
        "alphav": ((0, nx + 1), (-1, ny + 1)),
        "alphae": ((0, nx + 1), (0, ny + 1)),
        "isource_map": ((0, nx + 1), (0, ny + 1)),
        "alphaw": ((-1, nz + 1),),
    }
    for n in ("t_bkg", "s_bkg", "c_bkg", "rho_bkg", "azzmin", "pgradx", "pgrady"):
        table[n] = ((0, nz + 1),)
    return table.get(name, S)


def shape_of(b):
    return tuple(h - l + 1 for l, h in b)


class FArray:
    """A numpy array (Fortran order) with Fortran lower bounds."""

    def __init__(self, data, b):
        self.a = data
        self.b = b

    def __getitem__(self, idx):
        if not isinstance(idx, tuple):
            idx = (idx,)
        out = []
        for q, ix in enumerate(idx):
            l = self.b[q][0]
            if isinstance(ix, slice):
                out.append(slice(None if ix.start is None else ix.start - l,
                                 None if ix.stop is None else ix.stop - l + 1))
            else:
                out.append(ix - l)
        return self.a[tuple(out)]


STATE_FIELDS = ("u", "up", "um", "v", "vp", "vm", "w", "wp", "wm", "p", "t", "s",
                "c", "ss", "ss_v", "rhop", "fx", "fy", "fz")


# --------------------------------------------------------------------------
class Oracle:
    """ctypes wrapper of oracle/liboracle.so (the C restatement)."""
    _lib = None

    _path = ORACLE_SO

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = ctypes.CDLL(cls._path)
            L.o_create.restype = ctypes.c_void_p
            L.o_create.argtypes = [ctypes.POINTER(Params)]
            for f in ("o_destroy", "o_sponge_maker", "o_initialize", "o_body_force",
                      "o_forces", "o_pressure", "o_momentum_u", "o_momentum_v",
                      "o_momentum_w", "o_viscosity3d", "o_mixing", "o_work_velocity",
                      "o_density", "o_refresh"):
                getattr(L, f).argtypes = [ctypes.c_void_p]
                getattr(L, f).restype = None
            L.o_call.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
            L.o_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
            L.o_set.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
            L.o_bounds.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
            L.o_scalars.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            L.o_begin_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
            L.o_source_terms.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            L.o_tracer.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double]
            L.o_vent.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
            L.o_corrective.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int]
            L.o_get_tan.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
            L.o_set_tan.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
            L.o_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
            L.o_step.restype = ctypes.c_int
            L.o_cfl_values.argtypes = [ctypes.c_void_p]
            L.o_cfl_values.restype = ctypes.c_int
            L.o_init_profiles.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 3 + [ctypes.c_int]
            L.o_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
            L.o_hs3crt_c.argtypes = [ctypes.c_void_p] * 8
            L.o_rho_rho.argtypes = [ctypes.POINTER(Params), ctypes.c_double, ctypes.c_double]
            L.o_rho_rho.restype = ctypes.c_double
            L.o_source_patch.argtypes = [ctypes.POINTER(Params)] + [ctypes.c_void_p] * 4
            cls._lib = L
        return cls._lib

    def __init__(self, prm):
        self.prm = prm.copy()
        self.L = self.lib()
        self.h = self.L.o_create(ctypes.byref(self.prm))
        self.nx, self.ny, self.nz = prm.nx, prm.ny, prm.nz

    def __del__(self):
        try:
            self.L.o_destroy(self.h)
        except Exception:
            pass

    def bounds(self, name):
        return bounds(name, self.nx, self.ny, self.nz)

    def get(self, name):
        b = self.bounds(name)
        buf = np.zeros(shape_of(b), dtype=np.float64, order="F")
        if self.L.o_get(self.h, name.encode(), buf.ctypes.data) != 0:
            raise KeyError(name)
        return buf

    def getf(self, name):
        return FArray(self.get(name), self.bounds(name))

    def set(self, name, arr):
        b = self.bounds(name)
        buf = np.asfortranarray(np.asarray(arr, dtype=np.float64).reshape(shape_of(b), order="F"))
        if self.L.o_set(self.h, name.encode(), buf.ctypes.data) != 0:
            raise KeyError(name)

    def call(self, name):
        r = self.L.o_call(self.h, name.encode())
        if r < 0:
            raise KeyError(name)
        return r

    def begin_step(self, m):
        self.L.o_begin_step(self.h, m)

    def set_time(self, time, mtime):
        self.L.o_set_time(self.h, time, mtime)

    def source_terms(self, noise):
        noise = np.ascontiguousarray(noise, dtype=np.float64)
        self.L.o_source_terms(self.h, noise.ctypes.data)

    def tracer(self, name, cmin, cmax, tau3):
        self.L.o_tracer(self.h, name.encode(), cmin, cmax, tau3)

    def vent(self, name, value):
        self.L.o_vent(self.h, name.encode(), value)

    def step(self, m, noise):
        noise = np.ascontiguousarray(noise, dtype=np.float64)
        return self.L.o_step(self.h, m, noise.ctypes.data)

    def init_profiles(self, z, t, s):
        z, t, s = (np.ascontiguousarray(x, dtype=np.float64) for x in (z, t, s))
        self.L.o_init_profiles(self.h, z.ctypes.data, t.ctypes.data, s.ctypes.data, len(z))

    def scalars(self):
        out = np.zeros(32)
        self.L.o_scalars(self.h, out.ctypes.data)
        return out

    def patch(self):
        v = [ctypes.c_int() for _ in range(4)]
        self.L.o_source_patch(ctypes.byref(self.prm), *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)


class OracleCS(Oracle):
    """The oracle built with -DHC_CS (oracle/liboracle_cs.so): every state
    value carries a tangent by complex-step differentiation, so the oracle's
    routines are also their tangent-linear model.  get/set read and write
    the values, get_tan/set_tan the tangents (set_tan keeps the value)."""
    _lib = None
    _path = ORACLE_CS_SO


def _get_tan(self, name):
    b = self.bounds(name)
    buf = np.zeros(shape_of(b), dtype=np.float64, order="F")
    if self.L.o_get_tan(self.h, name.encode(), buf.ctypes.data) != 0:
        raise KeyError(name)
    return buf


def _set_tan(self, name, arr):
    b = self.bounds(name)
    buf = np.asfortranarray(np.asarray(arr, dtype=np.float64).reshape(shape_of(b), order="F"))
    if self.L.o_set_tan(self.h, name.encode(), buf.ctypes.data) != 0:
        raise KeyError(name)


# tangents of the state (only the complex-step build carries them; the plain
# build reads zeros and refuses to set one)
Oracle.get_tan = _get_tan
Oracle.set_tan = _set_tan


def rho_rho(prm, s, t):
    return Oracle.lib().o_rho_rho(ctypes.byref(prm), s, t)


# --------------------------------------------------------------------------
# profiles and noise
def linear_profile_table(zmin=1690.0, zmax=2110.0, dzt=2.0, t0=1.673, tg=6.666e-4,
                         s0=34.612, sg=-1.135e-4, zend=2101.0):
    """The analytic form of the reference's site_stratification_linear.inc
    (its header lines 1-3), sampled on its own grid."""
    z = np.arange(zmin, zmax + 0.5 * dzt, dzt)
    return z, t0 + tg * (zend - z), s0 + sg * (zend - z)


def read_profile_file(path):
    """unit-1010 table: 5 header lines then (3F15.7) rows."""
    z, t, s = [], [], []
    with open(path) as fh:
        lines = fh.read().splitlines()[5:]
    for ln in lines:
        if not ln.strip():
            continue
        z.append(float(ln[0:15])); t.append(float(ln[15:30])); s.append(float(ln[30:45]))
    return np.array(z), np.array(t), np.array(s)


def write_profile_file(path, z, t, s):
    with open(path, "w") as fh:
        for _ in range(5):
            fh.write(" header\n")
        for a, b, c in zip(z, t, s):
            fh.write("%15.7f%15.7f%15.7f\n" % (a, b, c))


def uniform_noise(seed, n):
    return np.random.default_rng(seed).random(3 * n)


# --------------------------------------------------------------------------
# compiled reference (oracle/_ref) in a fresh process
def ref_available():
    return os.path.isdir(REF_DIR) and shutil_which("amdflang")


def shutil_which(x):
    import shutil
    return shutil.which(x) is not None


def ref_lib_path(nx, ny, nz):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import build_ref
    path = build_ref.libpath(nx, ny, nz)
    srcs = [os.path.join(ROOT, "oracle", f) for f in ("ref_harness.f90", "build_ref.py")]
    if not os.path.exists(path) or os.path.getmtime(path) < max(os.path.getmtime(f) for f in srcs):
        build_ref.build(nx, ny, nz)
    return path


class RefLib:
    """The reference's subroutines through oracle/ref_harness.f90.  Module
    state is per process: use one instance per fresh process."""

    def __init__(self, path, prm, threads=1):
        self.L = ctypes.CDLL(path, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        # one thread unless asked: the reference's VISCOSITY3D shares its
        # strain scalars between threads (a race, see oracle/ref_harness.f90)
        self.L.href_set_threads.argtypes = [ctypes.c_int]
        self.L.href_set_threads(threads)
        self.prm = prm
        sz = (ctypes.c_int * 3)()
        self.L.href_sizes(sz)
        assert tuple(sz) == (prm.nx, prm.ny, prm.nz)
        self.nx, self.ny, self.nz = prm.nx, prm.ny, prm.nz
        self.L.href_copy.restype = ctypes.c_int
        self.L.href_copy.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        self.L.href_get_noise.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.L.href_scalars.argtypes = [ctypes.c_void_p]
        self.L.href_init_profiles.argtypes = [ctypes.c_char_p, ctypes.c_int]
        for f in ("href_step", "href_begin_step", "href_tracer"):
            getattr(self.L, f).argtypes = [ctypes.c_int]
        ip, rp = ref_int_real(prm)
        self.L.href_setup(ip, rp)

    def get(self, name):
        b = bounds(name, self.nx, self.ny, self.nz)
        buf = np.zeros(shape_of(b), dtype=np.float64, order="F")
        n = self.L.href_copy(name.encode(), len(name), buf.ctypes.data, 0)
        if n < 0:
            raise KeyError(name)
        assert n == buf.size, (name, n, buf.size)
        return buf

    def set(self, name, arr):
        b = bounds(name, self.nx, self.ny, self.nz)
        buf = np.asfortranarray(np.asarray(arr, dtype=np.float64).reshape(shape_of(b), order="F"))
        n = self.L.href_copy(name.encode(), len(name), buf.ctypes.data, 1)
        assert n == buf.size

    def call(self, name, *args):
        getattr(self.L, "href_" + name)(*args)

    def init_profiles(self, path):
        b = path.encode()
        self.L.href_init_profiles(b, len(b))

    def scalars(self):
        out = np.zeros(16)
        self.L.href_scalars(out.ctypes.data)
        return out

    def noise(self):
        buf = np.zeros(3 * 4096)
        n = ctypes.c_int()
        self.L.href_get_noise(buf.ctypes.data, ctypes.byref(n))
        return buf[: 3 * n.value].copy()


def _ref_worker(args):
    fn, path, prm_bytes, kw = args
    prm = Params.from_buffer_copy(prm_bytes)
    with tempfile.TemporaryDirectory() as d:
        os.chdir(d)
        ref = RefLib(path, prm)
        return fn(ref, **kw)


def run_in_ref(fn, prm, **kw):
    """Run fn(ref, **kw) in a fresh process with the reference built for the
    grid of prm; fn must be a module-level function returning picklable data."""
    path = ref_lib_path(prm.nx, prm.ny, prm.nz)
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(1, mp_context=mp.get_context("fork")) as ex:
        return ex.submit(_ref_worker, (fn, path, bytes(prm), kw)).result()


def rel_err(a, b, mask=None):
    a = np.asarray(a); b = np.asarray(b)
    if mask is not None:
        a = a[mask]; b = b[mask]
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0
