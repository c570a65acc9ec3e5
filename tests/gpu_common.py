"""Helpers shared by the GPU tests: the HIP library's Python harness and
replays of the golden reference runs."""
import os
import sys

import numpy as np

import hc_testlib as T

sys.path.insert(0, os.path.join(T.ROOT, "plumemodel-hotcross_amd"))
import hotcross as H  # noqa: E402

UPLOAD = ("u", "v", "w", "up", "vp", "wp", "um", "vm", "wm", "p", "t", "s", "c", "ss", "ss_v", "rhop")


def to_hc(prm):
    return H.HcParams.from_buffer_copy(bytes(prm))


def context_from_state(prm, state, precision=H.FP64):
    ctx = H.Context(to_hc(prm), precision)
    ctx.set_profiles(state["t_bkg"], state["s_bkg"], state["rho_bkg"], state["azzmin"])
    ctx.set_sponge(state["alphau"], state["alphav"], state["alphae"], state["alphaw"])
    for n in UPLOAD:
        ctx.upload(n, state[n])
    return ctx


def compare(a, b):
    """max |a-b| / max |b| over the cells the reference holds finite and
    below the 1e34 sentinel; the sentinel cells themselves must agree."""
    fin = np.isfinite(b)
    msk = fin & (np.abs(b) < 1e30)
    if not np.array_equal(a[fin & ~msk], b[fin & ~msk]):
        return np.inf
    return T.rel_err(a, b, msk)


def compare_abs(a, b):
    """max |a-b| over the same cells as compare()."""
    fin = np.isfinite(b)
    msk = fin & (np.abs(b) < 1e30)
    if not np.array_equal(a[fin & ~msk], b[fin & ~msk]):
        return np.inf
    return float(np.abs(a[msk] - b[msk]).max()) if msk.any() else 0.0
