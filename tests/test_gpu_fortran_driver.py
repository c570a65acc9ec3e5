"""The Fortran host layer end to end: plumemodel-hotcross_amd/fortran's
driver calls the drop-in subroutines (the reference's names and call
sequence) over ISO_C_BINDING; its result is replayed by the oracle from the
driver's own initial state and random_number draws."""
import os
import subprocess

import numpy as np
import pytest

import hc_testlib as T

pytestmark = pytest.mark.gpu

EXE = os.path.join(T.ROOT, "plumemodel-hotcross_amd", "fortran", "hotcross_driver")
ORDER = ("u", "v", "w", "up", "vp", "wp", "um", "vm", "wm", "p", "t", "s", "c", "ss", "ss_v", "rhop")


def run_driver(tmp, nx, ny, nz, nsteps):
    case = os.path.join(tmp, "case.nml")
    with open(case, "w") as fh:
        fh.write("&case nx=%d, ny=%d, nz=%d, nsteps=%d /\n" % (nx, ny, nz, nsteps))
    out = os.path.join(tmp, "out%d.bin" % nsteps)
    r = subprocess.run([EXE, case, out], cwd=tmp, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    raw = np.fromfile(out, dtype=np.float64)
    fields, q = {}, 0
    for n in ORDER:
        b = T.bounds(n, nx, ny, nz)
        cnt = int(np.prod(T.shape_of(b)))
        fields[n] = raw[q:q + cnt].reshape(T.shape_of(b), order="F")
        q += cnt
    noise = raw[q:].reshape((nsteps, -1)) if nsteps else None
    return fields, noise


def test_fortran_driver_matches_oracle(tmp_path):
    nx, ny, nz, n = 16, 12, 10, 12
    print("fortran driver", nx, ny, nz, n, flush=True)
    init, _ = run_driver(str(tmp_path), nx, ny, nz, 0)
    final, noise = run_driver(str(tmp_path), nx, ny, nz, n)
    prm = T.reference_params(nx, ny, nz)
    o = T.Oracle(prm)
    z = 1690.0 + 2.0 * np.arange(220)
    o.init_profiles(z, 1.673 + 6.666e-4 * (2101.0 - z), 34.612 - 1.135e-4 * (2101.0 - z))
    o.call("sponge_maker")
    for k in ORDER:
        o.set(k, init[k])
    o.set_time(-prm.dt, 0)
    for m in range(n):
        o.step(m, noise[m])
    for k in ORDER:
        b = o.get(k)
        a = final[k]
        msk = np.abs(b) < 1e30
        assert np.array_equal(a[~msk], b[~msk]), k
        if k == "rhop":
            # one-ulp differences of rho from s**1.5 (see test_gpu_parity)
            err = float(np.abs(a[msk] - b[msk]).max())
            assert err <= 8 * np.spacing(1030.0), (k, err)
        else:
            err = T.rel_err(a, b, msk)
            assert err <= 1e-10, (k, err)
