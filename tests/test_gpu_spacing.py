"""Grid spacings: the kernels divide by dx, dy, dz as the reference does, and
for power-of-two spacings multiply by the exact reciprocal instead (the same
correctly rounded value).  Both forms are checked: a general spacing against
the oracle, and a power-of-two spacing against the division form forced
with HC_NO_P2=1 (bit-identical)."""
import os

import numpy as np
import pytest

import hc_testlib as T
import gpu_common as G
from test_oracle_golden import load_case

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(T.ROOT, "tests", "golden", "ref_16x12x10_50.npz")
RHO_ULPS = 2 * float(np.spacing(1030.0))


def _case(dx, dy, dz, dt):
    _, before, _, noise, _ = load_case(GOLDEN)
    prm = T.reference_params(16, 12, 10, dx=dx, dy=dy, dz=dz, dt=dt)
    return prm, before, noise


def test_general_spacing_matches_oracle():
    prm, before, noise = _case(0.8, 0.6, 0.7, 0.2)
    print("spacing 0.8 0.6 0.7, 4 steps", flush=True)
    o = T.Oracle(prm)
    for k, v in before.items():
        o.set(k, v)
    o.set_time(-prm.dt, 0)
    ctx = G.context_from_state(prm, before)
    for m in range(4):
        o.step(m, noise[m])
        ctx.driver_step(m, noise[m])
    for f in ("u", "v", "w", "p", "t", "s", "c", "ss", "ss_v", "rhop"):
        a, b = ctx.download(f), o.get(f)
        if f == "rhop":
            assert G.compare_abs(a, b) <= 4 * RHO_ULPS, f
        else:
            err = G.compare(a, b)
            assert err <= 1e-10, (f, err)


def test_pow2_reciprocal_equals_division():
    prm, before, noise = _case(0.5, 0.5, 0.5, 0.1)
    print("spacing 0.5, 4 steps, multiply vs divide", flush=True)
    out = []
    for no_p2 in ("0", "1"):
        os.environ["HC_NO_P2"] = no_p2
        try:
            ctx = G.context_from_state(prm, before)
        finally:
            del os.environ["HC_NO_P2"]
        for m in range(4):
            ctx.driver_step(m, noise[m])
        out.append({f: ctx.download(f) for f in ("u", "v", "w", "p", "t", "s", "c", "ss", "ss_v", "rhop")})
        ctx.close()
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f], equal_nan=True), f
