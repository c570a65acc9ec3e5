"""The C-ABI library loads and exports every entry point include/hotcross.h
declares; the host-side struct mirrors agree (no device needed)."""
import ctypes
import os
import re

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
