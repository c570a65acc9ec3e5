#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference itself.

The reference's Fortran is compiled in place by oracle/build_ref.py (this
needs /root/reference and amdflang) and driven through oracle/ref_harness.f90:
reference set-up (INITIALIZE_PROFILES on the reference's
site_stratification_linear.inc, DENSITY, INITIALIZE, SPONGE_MAKER), a seeded
smooth perturbation, then N steps of the reference time loop body
(hotcross80-revised2014.f:264-350).  Each fixture holds the full starting
state, the random_number draws SOURCE_terms made at every step, and the
reference's state after N steps; tests replay the start through the oracle
and the HIP library and compare.

usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import hc_testlib as T  # noqa: E402
import refcases as R  # noqa: E402

CASES = [
    # name, grid, steps, bit7, bcondy
    ("ref_16x12x10_50", (16, 12, 10), 50, 1, 1),
    ("ref_24x16x12_20_norot", (24, 16, 12), 20, 0, 1),
    ("ref_16x12x10_3_yper", (16, 12, 10), 3, 1, 0),
]


def main():
    for name, (nx, ny, nz), n, bit7, by in CASES:
        prm = T.reference_params(nx, ny, nz, bit7=bit7, bcondy=by)
        before, after, noises, _ = T.run_in_ref(R.worker_steps, prm, seed=7, nsteps=n)
        out = {"prm": np.frombuffer(bytes(prm), dtype=np.uint8), "nsteps": np.array(n),
               "noise": np.array(noises)}
        for k, v in before.items():
            out["before_" + k] = v
        for k in T.STATE_FIELDS:
            out["after_" + k] = after[k]
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
