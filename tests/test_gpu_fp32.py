"""The FP32 chosen form against the FP64 oracle (error stated per field).

FP32 storage and arithmetic for the stencils, MPDATA and the x/y transforms;
FP64 for the equation of state and Richardson number, the singular (0,0)
pressure mode, the RHS total and the budgets."""
import glob
import os

import numpy as np
import pytest

import hc_testlib as T
import gpu_common as G
from test_oracle_golden import load_case

pytestmark = pytest.mark.gpu

# relative max error per field after the 50-step golden run (vs max |field|).
# Measured on MI355X: u 4.3e-5, v 1.65e-4, w 1.04e-3, p 4.5e-4, t 5.7e-6,
# s 4.9e-6, c 7.4e-6, rhop 2.8e-3; the bounds are about twice those.
TOL = {"u": 1e-4, "v": 3e-4, "w": 2e-3, "p": 1e-3, "t": 1e-5, "s": 1e-5, "c": 2e-5, "rhop": 5e-3}


def test_fp32_form_error_vs_fp64_reference():
    path = os.path.join(T.ROOT, "tests", "golden", "ref_16x12x10_50.npz")
    prm, before, after, noise, n = load_case(path)
    ctx = G.context_from_state(prm, before, precision=G.H.FP32)
    for m in range(n):
        ctx.driver_step(m, noise[m])
    errs = {}
    for k, tol in TOL.items():
        b = after[k]
        msk = np.abs(b) < 1e30
        errs[k] = T.rel_err(ctx.download(k), b, msk)
    print("fp32 errors vs reference after %d steps:" % n, {k: "%.2e" % v for k, v in errs.items()})
    for k, tol in TOL.items():
        assert errs[k] <= tol, (k, errs[k])
