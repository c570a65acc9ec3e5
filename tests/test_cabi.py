"""The C-ABI library loads and exports every entry point include/hotcross.h
declares; the host-side struct mirrors agree (no device needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

import hc_testlib as T
import gpu_common as G

HEADER = os.path.join(T.ROOT, "include", "hotcross.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hc_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("hc_create", "hc_forces", "hc_pressure", "hc_tracer31", "hc_step", "hc_refresh"):
        assert n in names


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(G.H.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_params_mirrors_agree():
    assert ctypes.sizeof(G.H.HcParams) == ctypes.sizeof(T.Params)
    assert [f[0] for f in G.H.HcParams._fields_] == [f[0] for f in T.Params._fields_]
    # layout of the C struct: the offsets the C compiler gives
    assert G.H.HcParams.dx.offset == 16 and G.H.HcParams.cs_h.offset % 8 == 0


def test_create_rejects_bad_grid_without_device_work():
    p = G.to_hc(T.reference_params(16, 12, 10))
    p.nx = 10  # 2*5: not a product of 2 and 3
    h = ctypes.c_void_p()
    r = G.H.lib().hc_create(ctypes.byref(p), 0, ctypes.byref(h))
    assert r == -1
    assert b"grid" in G.H.lib().hc_last_error(h)
    G.H.lib().hc_destroy(h)


def stat_sums(samples):
    """Running sums of hc_stats_get for a list of per-sample dicts of
    (wpoint, t, uc, vc, wc) values on a one-cell slab."""
    s = np.zeros((G.H.NSTAT, 1, 1, 1))
    for x in samples:
        w, t, u, v, c = x["w"], x["t"], x["u"], x["v"], x["wc"]
        s[:, 0, 0, 0] += (w, t, u, v, c, u * u, v * v, c * c, t * t, c * t)
    return s, len(samples)


def test_plane_statistics_closed_forms():
    # two samples: u = 1 +- 2, v = 0, cell w = 1 and 3, T = 2 and 6, w point = 1.5 and 2.5
    st = G.H.plane_statistics(*stat_sums([
        {"w": 1.5, "t": 2.0, "u": -1.0, "v": 0.0, "wc": 1.0},
        {"w": 2.5, "t": 6.0, "u": 3.0, "v": 0.0, "wc": 3.0}]))
    # w is stored positive downward; the statistics report it upward
    assert st["w_up"][0, 0, 0] == -2.0
    assert st["t"][0, 0, 0] == 4.0
    # TKE = 0.5 * (var u + var w) = 0.5 * (4 + 1)
    assert st["tke"][0, 0, 0] == 2.5
    assert st["t_var"][0, 0, 0] == 4.0
    # cov(w_cell, T) = ((1-2)(2-4) + (3-2)(6-4)) / 2 = 2, reported with w upward
    assert st["wt_flux"][0, 0, 0] == -2.0
    # a steady state has no fluctuations
    st = G.H.plane_statistics(*stat_sums([{"w": 0.3, "t": 1.7, "u": 0.1, "v": -0.2, "wc": 0.3}] * 5))
    assert abs(st["tke"][0, 0, 0]) < 1e-15 and abs(st["wt_flux"][0, 0, 0]) < 1e-15
    with pytest.raises(ValueError):
        G.H.plane_statistics(np.zeros((G.H.NSTAT, 1, 1, 1)), 0)
