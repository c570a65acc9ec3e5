"""The netCDF writer (ncwrite.py) read back by an independent reader
(scipy.io.netcdf_file), and the writers' host-side pieces (no device)."""
import os
import sys

import numpy as np
import pytest
from scipy.io import netcdf_file

import hc_testlib as T

sys.path.insert(0, os.path.join(T.ROOT, "plumemodel-hotcross_amd"))
import ncwrite  # noqa: E402
import output  # noqa: E402


def test_records_and_fixed_variables_read_back(tmp_path):
    rng = np.random.default_rng(0)
    nx, ny, nz = 5, 4, 3
    path = str(tmp_path / "t.nc")
    recs = [rng.random((nx + 1, ny, nz)).astype(np.float32) for _ in range(3)]
    c = rng.random((nx, ny, nz)).astype(np.float32)
    with ncwrite.NcWriter(path, [("time", None), ("x", nx), ("ux", nx + 1), ("y", ny), ("z", nz)],
                          [("time", "f8", ("time",), {"units": "sec"}),
                           ("x", "f4", ("x",), {"units": "m", "positive": "down"}),
                           ("u", "f4", ("time", "z", "y", "ux"), {"missing_value": np.float32(-1e34)}),
                           ("c", "f4", ("z", "y", "x"), {"long_name": "Tracer-0"})],
                          {"title": "test".ljust(80)}) as w:
        w.put("x", np.arange(nx, dtype=np.float32) * 0.5)
        w.put("c", c)
        for r, u in enumerate(recs):
            w.append({"time": np.array([1.5 * r]), "u": u})
    f = netcdf_file(path, "r", mmap=False)
    try:
        assert f.version_byte == 2
        assert f.dimensions == {"time": None, "x": nx, "ux": nx + 1, "y": ny, "z": nz}
        assert f.title == b"test".ljust(80)
        assert f.variables["x"].positive == b"down"
        assert np.array_equal(f.variables["time"][:], [0.0, 1.5, 3.0])
        # Fortran arrays (x first) appear with the dimensions reversed
        for r in range(3):
            assert np.array_equal(f.variables["u"][r], recs[r].transpose(2, 1, 0))
        assert np.array_equal(f.variables["c"][:], c.transpose(2, 1, 0))
        assert f.variables["u"].missing_value == np.float32(-1e34)
    finally:
        f.close()


def test_lone_record_variable_and_odd_sizes(tmp_path):
    path = str(tmp_path / "o.nc")
    with ncwrite.NcWriter(path, [("t", None), ("n", 3)],
                          [("a", "i2", ("t", "n"), {}), ("b", "i2", ("n",), {"scale": 2.5})]) as w:
        w.put("b", np.array([1, 2, 3], dtype=np.int16))
        w.append({"a": np.array([4, 5, 6], dtype=np.int16)})
        w.append({"a": np.array([7, 8, 9], dtype=np.int16)})
    f = netcdf_file(path, "r", mmap=False)
    try:
        assert np.array_equal(f.variables["a"][:], [[4, 5, 6], [7, 8, 9]])
        assert np.array_equal(f.variables["b"][:], [1, 2, 3])
        assert f.variables["b"].scale == 2.5
    finally:
        f.close()


def test_writer_rejects_bad_definitions(tmp_path):
    with pytest.raises(ValueError):
        ncwrite.NcWriter(str(tmp_path / "a.nc"), [("t", None), ("u", None)], [])
    with pytest.raises(ValueError):
        ncwrite.NcWriter(str(tmp_path / "b.nc"), [("t", None), ("n", 2)], [("a", "f4", ("n", "t"), {})])
    w = ncwrite.NcWriter(str(tmp_path / "c.nc"), [("t", None), ("n", 2)], [("a", "f4", ("t", "n"), {})])
    with pytest.raises(ValueError):
        w.append({"a": np.zeros(3, dtype=np.float32)})
    w.close()


def test_fill_edges_counts():
    # whole box a x b x c: 4(a-2) + 4(b-2) + 4(c-2) + 8 edge cells
    a = np.zeros((6, 5, 7), dtype=np.float32, order="F")
    output.fill_edges(a, (0, 0, -1), (0, 0, -1), (5, 4, 5))
    assert (a == output.MISSING).sum() == 4 * (6 + 5 + 7) - 16
    assert a[0, 0, 3] == output.MISSING and a[2, 0, -1] == output.MISSING and a[2, 2, 3] == 0
    # a slab j = 2..3 strictly inside in y: only the four x-z edges of each row
    s = np.zeros((6, 2, 7), dtype=np.float32, order="F")
    output.fill_edges(s, (0, 2, -1), (0, 0, -1), (5, 4, 5))
    assert (s == output.MISSING).sum() == 4 * 2
    assert s[0, 1, 0] == output.MISSING and s[5, 0, 6] == output.MISSING and s[0, 1, 3] == 0
