"""hc_cfl_values_noalloc (CFL_VALUES over caller-owned device buffers, its
partial results in a static device array) against hc_cfl_values on a
context holding the same velocities: the same six Courant numbers bit for
bit and the same bailout codes, on two calls in a row, the second on a
smaller grid, in both precisions."""
import ctypes

import numpy as np
import pytest

import hc_testlib as T
import gpu_common as G

pytestmark = pytest.mark.gpu
H = G.H


def _lib():
    L = H.lib()
    f = L.hc_cfl_values_noalloc
    f.argtypes = [ctypes.c_int32] * 3 + [ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_double] * 4 + \
        [ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f


def _padded(arr, name, nx, ny, nz, dtype):
    """A Fortran-bounded host field in the library's device layout."""
    px = 16 * ((nx + 18 + 15) // 16)
    pad = np.zeros((nz + 3, ny + 3, px), dtype=dtype)
    (lo, hi) = H.field_bounds(name, nx, ny, nz)
    a = np.asarray(arr).reshape(tuple(h - l + 1 for l, h in zip(lo, hi)), order="F")
    pad[lo[2] + 1:hi[2] + 2, lo[1] + 1:hi[1] + 2, lo[0] + 16:hi[0] + 17] = np.transpose(a, (2, 1, 0))
    return pad.ravel()


def _case(nx, ny, nz, seed, courant):
    p = G.to_hc(T.reference_params(nx, ny, nz))
    rng = np.random.default_rng(seed)
    vel = {}
    for name, d in (("u", p.dx), ("v", p.dy), ("w", p.dz)):
        lo, hi = H.field_bounds(name, nx, ny, nz)
        shape = tuple(h - l + 1 for l, h in zip(lo, hi))
        vel[name] = np.asfortranarray(courant * d / p.dt * (2.0 * rng.random(shape) - 1.0))
    return p, vel


def _reference(p, vel, prec):
    ctx = H.Context(p, prec)
    try:
        for n, a in vel.items():
            ctx.upload(n, a)
        try:
            return ctx.cfl_values(), 0
        except H.HcError as e:
            return None, e.code
    finally:
        ctx.close()


def _hip():
    """hipMalloc, hipMemcpy and hipFree of the HIP runtime the library is
    bound to, looked up through the library's own handle (whose scope holds
    its dependencies): the caller's buffers must come from that runtime, and
    torch, which bundles another, cannot provide them in this process."""
    hip = ctypes.CDLL(H.LIB_PATH)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    for fn in (hip.hipMalloc, hip.hipMemcpy, hip.hipFree):
        fn.restype = ctypes.c_int
    return hip


def _noalloc(f, p, vel, prec):
    hip = _hip()
    dt = np.float64 if prec == H.FP64 else np.float32
    bufs = []
    try:
        for n in ("u", "v", "w"):
            a = _padded(vel[n], n, p.nx, p.ny, p.nz, dt)
            d = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(d), a.nbytes) == 0
            bufs.append(d)
            assert hip.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        out = np.zeros(6)
        r = f(p.nx, p.ny, p.nz, prec, *bufs, p.dx, p.dy, p.dz, p.dt, None, out.ctypes.data)
        return out, r
    finally:
        for d in bufs:
            hip.hipFree(d)


@pytest.mark.parametrize("prec", [H.FP64, H.FP32], ids=["fp64", "fp32"])
def test_noalloc_cfl_matches_context_entry(prec):
    f = _lib()
    # two calls in a row, the second on a smaller grid; then a state over
    # the 0.75 bailout on the small grid
    for (nx, ny, nz), seed, courant in (((48, 24, 20), 1, 0.3), ((24, 16, 12), 2, 0.5), ((24, 16, 12), 3, 0.9)):
        print("noalloc cfl", prec, nx, ny, nz, courant, flush=True)
        p, vel = _case(nx, ny, nz, seed, courant)
        ref, code = _reference(p, vel, prec)
        out, r = _noalloc(f, p, vel, prec)
        assert r == code, (r, code)
        if code == 0:
            assert np.array_equal(out, ref), (out, ref)
            assert np.abs(out).max() > 0.5 * courant
    # a non-finite velocity: HC_ENAN and NaN results, as hc_cfl_values
    p, vel = _case(24, 16, 12, 4, 0.2)
    vel["v"][3, 4, 5] = np.nan
    out, r = _noalloc(f, p, vel, prec)
    assert r == H.HC_ENAN and np.isnan(out).all()


def test_noalloc_cfl_rejects_bad_arguments():
    f = _lib()
    out = np.zeros(6)
    dummy = ctypes.c_void_p(16)  # never dereferenced: the arguments are refused first
    assert f(24, 16, 12, 7, dummy, dummy, dummy, 1.0, 1.0, 1.0, 0.25, None, out.ctypes.data) == H.HC_EINVAL
    assert f(0, 16, 12, H.FP64, dummy, dummy, dummy, 1.0, 1.0, 1.0, 0.25, None, out.ctypes.data) == H.HC_EINVAL
    assert f(24, 16, 12, H.FP64, None, dummy, dummy, 1.0, 1.0, 1.0, 0.25, None, out.ctypes.data) == H.HC_EINVAL
    # a padded box of 2^31 elements or more
    assert f(4096, 4096, 200, H.FP64, dummy, dummy, dummy, 1.0, 1.0, 1.0, 0.25, None,
             out.ctypes.data) == H.HC_EINVAL
