"""The tiled MPDATA corrective pass (the phases of k_mpdata_pass, hc_mpdata.h,
of the fixed-assignment tile, hc_mpdata2.h, and of the tile that forms each
face's donor-cell flux once, hc_mpdata3.h), built for the host
(tests/mpcheck) and run block by block, against one corrective pass of the
oracle's TRACER31 (mpdata-hotcross80.f:196-381), bit for bit in FP64.

The grid (72 x 36 x 40) is not a multiple of any tile (32 or 16 wide, 16 or
8 rows), and the planes split into chunks of 2, 4, 8 and 40, so every x, y and
z seam of the tiles, the partial tiles at the far x and y ends and the
domain's first faces (ud at i = 0, vd at j = 0, wd at k = 0) meet the
oracle.  LDS starts each block filled with NaN and every LDS index is
range-checked, so a read of a slot that no phase wrote, or an index outside
its plane, fails the test.  The limiter bounds sit at the 10th and 90th
percentiles of the field, so both flags occur in about a fifth of the cells."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import hc_testlib as T

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "mpcheck", "mpcheck.hip")
LIB = os.path.join(HERE, "mpcheck", "libmpcheck.so")
CSRC = os.path.join(T.ROOT, "plumemodel-hotcross_amd", "csrc")
NX, NY, NZ = 72, 36, 40


def build():
    deps = [SRC] + [os.path.join(CSRC, f) for f in ("hc_mpdata.h", "hc_mpdata2.h", "hc_mpdata3.h", "hc_common.h")]
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(f) for f in deps):
        subprocess.run(["hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
                        "-ffp-contract=off", "-Wno-unused-result", "-o", LIB, SRC], check=True,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    L = ctypes.CDLL(LIB)
    L.mpc_layout.restype = ctypes.c_int64
    L.mpc_pass.restype = ctypes.c_int64
    L.mpc_pass.argtypes = [ctypes.c_int] * 6 + [ctypes.c_void_p] * 15
    return L


@pytest.fixture(scope="module")
def lib():
    return build()


def layout(nx, ny, nz):
    return 16 * ((nx + 18 + 15) // 16), ny + 3, nz + 3


def pad(a, name, nx, ny, nz):
    """A Fortran-bounded field in the library's padded layout (zeros
    elsewhere)."""
    px, py, pz = layout(nx, ny, nz)
    out = np.zeros((pz, py, px))
    b = T.bounds(name, nx, ny, nz)
    out[b[2][0] + 1:b[2][1] + 2, b[1][0] + 1:b[1][1] + 2, b[0][0] + 16:b[0][1] + 17] = np.transpose(a, (2, 1, 0))
    return out.ravel()


def unpad(flat, name, nx, ny, nz):
    px, py, pz = layout(nx, ny, nz)
    v = flat.reshape(pz, py, px)
    b = T.bounds(name, nx, ny, nz)
    return np.transpose(v[b[2][0] + 1:b[2][1] + 2, b[1][0] + 1:b[1][1] + 2, b[0][0] + 16:b[0][1] + 17], (2, 1, 0))


def setup(seed):
    p = T.reference_params(NX, NY, NZ)
    o = T.Oracle(p)
    o.call("sponge_maker")
    rng = np.random.default_rng(seed)
    c = 1.0 + 0.5 * rng.random(T.shape_of(o.bounds("t")))
    vel = {}
    for f in ("uc", "vc", "wc"):
        vel[f] = 0.3 * (2.0 * rng.random(T.shape_of(o.bounds(f))) - 1.0)
    return p, o, c, vel


def host_pass(L, variant, tile, kc, c, vel, cmin, cmax, last=0, entry=None, tau3=0.25, cbkg=None, q=None, o=None):
    n = L.mpc_layout(NX, NY, NZ)
    arr = lambda a, f: np.ascontiguousarray(pad(a, f, NX, NY, NZ))
    cc = arr(c, "t")
    U, V, W = arr(vel["uc"], "uc"), arr(vel["vc"], "vc"), arr(vel["wc"], "wc")
    ud, vd, wd = (np.zeros(n) for _ in range(3))
    cout = arr(entry, "t") if last else np.zeros(n)
    cpre = np.zeros(n)
    dbl = np.array([cmin, cmax, tau3, 1.0])
    ints = np.zeros(6, dtype=np.int32)
    extra = [None] * 4
    if last:
        ext = o.getf("alphau"), o.getf("alphav")
        au, av = ext
        alphae0 = np.zeros((NX + 2, NY + 2))
        for j in range(NY + 2):  # alphae before its periodic copies, sponge80.f:107-112
            for i in range(NX + 2):
                alphae0[i, j] = max(au[i, j] + au[i - 1, j], av[i, j] + av[i, j - 1]) / 2.0
        nxs, nxe, nys, nye = 30, 33, 15, 19
        qpatch = np.array([q[i, j, NZ] for j in range(nys, nye + 1) for i in range(nxs, nxe + 1)])
        ints[:] = [1, 1, nxs, nxe, nys, nye]
        extra = [np.ascontiguousarray(cbkg), np.ascontiguousarray(alphae0.ravel(order="F")),
                 np.ascontiguousarray(o.get("alphaw")), np.ascontiguousarray(qpatch)]
    ptr = lambda a: None if a is None else a.ctypes.data
    bad = L.mpc_pass(variant, tile, NX, NY, NZ, kc, *[ptr(x) for x in (cc, U, V, W, ud, vd, wd, cout, cpre, dbl,
                                                                      ints)], *[ptr(x) for x in extra])
    assert bad == 0, "LDS or field index out of range %d times" % bad
    return {"c": unpad(cout, "t", NX, NY, NZ), "cpre": unpad(cpre, "t", NX, NY, NZ),
            "ud": unpad(ud, "uc", NX, NY, NZ), "vd": unpad(vd, "vc", NX, NY, NZ), "wd": unpad(wd, "wc", NX, NY, NZ)}


VARIANTS = [(1, 0, 4), (1, 0, 8), (1, 2, 40), (1, 1, 8), (2, 0, 4), (2, 0, 8), (2, 0, 40), (2, 1, 4),
            (3, 0, 2), (3, 0, 4), (3, 0, 8), (3, 0, 40), (3, 1, 4)]
IDS = ["%s_%s_kc%d" % (("mptile", "mptile2", "mptile3")[v - 1], ("32x16", "16x16", "32x8")[t], k)
       for v, t, k in VARIANTS]


@pytest.mark.parametrize("variant,tile,kc", VARIANTS, ids=IDS)
def test_corrective_pass_matches_oracle(lib, variant, tile, kc):
    p, o, c, vel = setup(1)
    core = c[1:NX + 1, 1:NY + 1, 1:NZ + 1]
    cmin, cmax = np.quantile(core, 0.1), np.quantile(core, 0.9)
    for f, a in vel.items():
        o.set(f, a)
    o.set("t", c)
    assert o.L.o_corrective(o.h, b"t", 2, cmin, cmax, 0.25, 0) == 0
    ref_c = o.get("t")
    ref_ud, ref_vd, ref_wd = o.get("uc"), o.get("vc"), o.get("wc")  # the pass copies ud, vd, wd there
    got = host_pass(lib, variant, tile, kc, c, vel, cmin, cmax)
    s = (slice(1, NX + 1), slice(1, NY + 1), slice(1, NZ + 1))
    assert np.array_equal(got["c"][s], ref_c[s])
    # the flags fired: the limited step differs from the unlimited one
    assert np.count_nonzero(ref_c[s] <= cmin) + np.count_nonzero(ref_c[s] >= cmax) > 0
    # faces the pass writes: ud over i = 0..nx (u-bounds start at -1), vd over
    # j = 0..ny, wd over k = 0..nz
    assert np.array_equal(got["ud"][1:NX + 2, 1:NY + 1, 1:NZ + 1], ref_ud[1:NX + 2, 1:NY + 1, 1:NZ + 1])
    assert np.array_equal(got["vd"][1:NX + 1, 1:NY + 2, 1:NZ + 1], ref_vd[1:NX + 1, 1:NY + 2, 1:NZ + 1])
    assert np.array_equal(got["wd"][1:NX + 1, 1:NY + 1, 1:NZ + 2], ref_wd[1:NX + 1, 1:NY + 1, 1:NZ + 2])


@pytest.mark.parametrize("variant,tile,kc", [(1, 0, 8), (2, 0, 4), (2, 1, 8), (3, 0, 4), (3, 1, 8)],
                         ids=["mptile_32x16_kc8", "mptile2_32x16_kc4", "mptile2_16x16_kc8", "mptile3_32x16_kc4",
                              "mptile3_16x16_kc8"])
def test_last_corrective_pass_with_sponge_and_source(lib, variant, tile, kc):
    """The last pass (iord = 3) with the sponge relative to the entry value
    and the vent source on its patch at k = nz (mpdata-hotcross80.f:356-381);
    the cells next to a boundary also return their pre-sponge value, which
    the boundary images take."""
    p, o, c, vel = setup(2)
    rng = np.random.default_rng(9)
    entry = c + 0.01 * rng.random(c.shape)
    cbkg = 1.2 + 0.01 * np.arange(NZ + 2)
    q = np.zeros(T.shape_of(o.bounds("q_t")))
    q[30:34, 15:20, NZ] = 2.0 + rng.random((4, 5))
    core = c[1:NX + 1, 1:NY + 1, 1:NZ + 1]
    cmin, cmax = np.quantile(core, 0.1), np.quantile(core, 0.9)
    for f, a in vel.items():
        o.set(f, a)
    o.set("t", c)
    o.set("cp", entry)
    o.set("q_t", q)
    o.set("t_bkg", cbkg)
    assert o.L.o_corrective(o.h, b"t", 3, cmin, cmax, 0.25, 0) == 0
    pre = o.get("t")
    for f, a in vel.items():  # the pass left its ud, vd, wd in uc, vc, wc
        o.set(f, a)
    o.set("t", c)
    assert o.L.o_corrective(o.h, b"t", 3, cmin, cmax, 0.25, 1) == 0
    ref = o.get("t")
    got = host_pass(lib, variant, tile, kc, c, vel, cmin, cmax, last=1, entry=entry, cbkg=cbkg, q=q, o=o)
    s = (slice(1, NX + 1), slice(1, NY + 1), slice(1, NZ + 1))
    assert np.array_equal(got["c"][s], ref[s])
    edge = np.zeros(c.shape, dtype=bool)
    edge[[1, NX], 1:NY + 1, 1:NZ + 1] = True
    edge[1:NX + 1, [1, NY], 1:NZ + 1] = True
    edge[1:NX + 1, 1:NY + 1, [1, NZ]] = True
    assert np.array_equal(got["cpre"][edge], pre[edge])
    assert not np.array_equal(ref[s], pre[s])  # the sponge and source acted
