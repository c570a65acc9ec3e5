"""Case set-up from files (case.py): the unit-1010 stratification reader
against the reference's own Fortran read of the same table, Fortran input
edit rules, the case-file reader and the reference's switch checks."""
import os
import shutil
import sys

import numpy as np
import pytest

import hc_testlib as T
import refcases as R

sys.path.insert(0, os.path.join(T.ROOT, "plumemodel-hotcross_amd"))
import case  # noqa: E402
import hotcross as H  # noqa: E402

HAVE_REF = os.path.isdir(T.REF_DIR) and shutil.which("amdflang")


def odd_table(path):
    """A table whose rows use the F15.7 input forms a Fortran runtime
    accepts: plain, implied decimals (no point), D and bare-sign exponents,
    embedded blanks, short rows padded with blanks."""
    z = np.arange(1700.0, 2102.0, 2.0)
    t = 1.673 + 6.666e-4 * (2101.0 - z)
    s = 34.612 - 1.135e-4 * (2101.0 - z)
    rows = []
    for q, (a, b, c) in enumerate(zip(z, t, s)):
        form = q % 5
        if form == 0:
            rows.append("%15.7f%15.7f%15.7f" % (a, b, c))
        elif form == 1:   # no decimal point: 7 implied decimals
            rows.append("%15.7f%15d%15d" % (a, round(b * 1e7), round(c * 1e7)))
        elif form == 2:   # D exponent and a bare-sign exponent
            rows.append("%15.7f%15s%15s" % (a, ("%.7E" % b).replace("E", "D"), "%.6f+1" % (c / 10)))
        elif form == 3:   # blanks inside the fields
            bs = "%.7f" % b
            rows.append("%15.7f%15s%15s" % (a, bs[:3] + "  " + bs[3:], " ".join("%.5f" % c)))
        else:             # lower-case e, short row padded
            rows.append(("%15.7f%15s%15s" % (a, "%.7e" % b, "%.7f" % c)).rstrip())
    with open(path, "w") as fh:
        fh.write("header 1\nheader 2\nheader 3\nheader 4\nheader 5\n")
        fh.write("\n".join(rows) + "\n")
    return z


def test_fortran_input_edits():
    f = case._fortran_f
    assert f("   1826.0000000") == 1826.0
    assert f("      18706784") == 1.8706784        # implied decimals
    assert f(" 1.5+3") == 1500.0 and f("-2.5D-1") == -0.25 and f("-12E2") == -12e-5
    assert f("") == 0.0 and f("     ") == 0.0 and f("1 2 3") == 123e-7 and f(" 7.") == 7.0
    for bad in ("abc", "1.2.3", "--1", "."):
        with pytest.raises(case.CaseError):
            f(bad)


@pytest.mark.skipif(not HAVE_REF, reason="reference sources / amdflang not present")
def test_table_reads_as_the_reference_reads_it(tmp_path):
    path = str(tmp_path / "odd.inc")
    odd_table(path)
    prm = T.reference_params(16, 12, 10)
    t_ref, s_ref = T.run_in_ref(R.worker_profiles, prm, path=path)
    z, t, s = case.read_stratification(path)
    assert len(z) == 201
    o = T.Oracle(prm)
    o.init_profiles(z, t, s)
    assert np.array_equal(o.get("t_bkg"), t_ref)
    assert np.array_equal(o.get("s_bkg"), s_ref)


def test_table_errors(tmp_path):
    p = tmp_path / "bad.inc"
    p.write_text("h\nh\nh\nh\nh\n%15.7f%15s%15.7f\n" % (1.0, "x1", 3.0))
    with pytest.raises(case.CaseError, match="error reading site_stratification_data.inc"):
        case.read_stratification(str(p))
    p.write_text("h\nh\n")
    with pytest.raises(case.CaseError):
        case.read_stratification(str(p))
    # the reference's read loop stops after 3000 rows without reaching the end
    z = np.arange(3000.0)
    case.write_stratification(str(p), z, z, z)
    with pytest.raises(case.CaseError):
        case.read_stratification(str(p))
    case.write_stratification(str(p), z[:2999], z[:2999], z[:2999])
    assert len(case.read_stratification(str(p))[0]) == 2999


def test_case_file_round_trip_and_reference_forms(tmp_path):
    p = H.workload_params("les128")
    p.cs_h, p.dt = 0.17, 0.125
    p.bdbit[2] = 0
    path = str(tmp_path / "run.case")
    case.write_case(p, path, {"title": "Dante Exp 1"})
    q, extras = case.read_case(path, base=H.workload_params("les64"))
    assert bytes(q) == bytes(p) and extras == {"title": "Dante Exp 1"}
    # forms of the reference's Fortran source
    src = ("c     a fixed-form comment line\n"
           "      nx = 96; ny = 64   ! two statements\n"
           "      dt = 2.5d-1\n"
           "      toggle(6) = 0\n"
           "      bdbit(1:2) = 0\n"
           "      wee = 0.025D0\n"
           "      title = 'Dante; Exp 1'\n"
           "      output_file_2 = \"snap.nc\"\n"
           "      open_box = .true.\n")
    (tmp_path / "ref.case").write_text(src)
    q, extras = case.read_case(str(tmp_path / "ref.case"), base=H.workload_params("les64"))
    assert (q.nx, q.ny, q.dt, q.toggle_t, q.noise_amp, q.open_box) == (96, 64, 0.25, 0, 0.025, 1)
    assert list(q.bdbit) == [0, 0, 1, 1]
    assert extras == {"title": "Dante; Exp 1", "output_file_2": "snap.nc"}


@pytest.mark.parametrize("stmt,msg", [
    ("bcondx = 2", "bcondx must be either 1"), ("bcondy = -1", "bcondy must be either 1"),
    ("bottom_slip = 0", "bottom_slip must be either -1"), ("side_slip = 2", "side_slip must be either -1")])
def test_case_checks_of_the_reference(tmp_path, stmt, msg):
    path = tmp_path / "c.case"
    path.write_text("      " + stmt + "\n")
    with pytest.raises(case.CaseError, match=msg):
        case.read_case(str(path))
