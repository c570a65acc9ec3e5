"""The oracle's tangent-linear build (oracle/liboracle_cs.so: the same source
compiled with -DHC_CS, complex-step differentiation) against the plain
oracle.  Its values are the plain oracle's up to roundoff (complex division
rounds its real part differently); its tangents are the derivative of the
plain oracle's map along the perturbation: central differences of the plain
oracle agree to their own truncation error, at a step whose limiter flags
and |x| kinks stay put, and the tangent is linear in the direction to
roundoff, which a difference quotient is not.  The direction leaves the
exact zeros of the state (the walls) alone: |x| has no derivative at 0,
where the tangent model, like the adjoint, takes 0.  The golden start state
holds v and w of 1e-19 on the lid and wall faces (roundoff of the
projection); the case sets them to the zeros they stand for, since any
difference step straddles |x| there while the tangent takes the slope on
the side of the roundoff's sign."""
import glob
import os

import numpy as np
import pytest

import hc_testlib as T
from test_oracle_golden import load_case

GOLDEN = [p for p in sorted(glob.glob(os.path.join(T.ROOT, "tests", "golden", "*.npz"))) if "16x12x10_50" in p][0]
STATE = ("u", "v", "w", "um", "vm", "wm", "up", "vp", "wp", "t", "s", "c", "ss", "ss_v", "rhop")
OUT = ("u", "v", "w", "p", "t", "s", "c", "ss", "ss_v", "rhop")
NSTEPS = 3


def _start(cls, before):
    prm = before[0]
    o = cls(prm)
    for k, v in before[1].items():
        o.set(k, v)
    o.set_time(-prm.dt, 0)
    return o


def _direction(seed, state):
    rng = np.random.default_rng(seed)
    d = {}
    for f in STATE:
        a = state[f]
        scale = {"ss": 1e-5, "ss_v": 1e-5, "t": 1e-3, "s": 1e-4, "c": 1e-2, "rhop": 1e-3}.get(f, 1e-3)
        x = scale * rng.standard_normal(a.shape)
        # sentinels stay put, and so do the exact zeros (walls): |x| at
        # x = 0 has no derivative, and the model (like its adjoint) takes 0
        x[~(np.abs(a) < 1e30) | (a == 0.0)] = 0.0
        d[f] = x
    return d


@pytest.fixture(scope="module")
def case():
    prm, before, _, noise, _ = load_case(GOLDEN)
    st = {}
    for f, a in before.items():
        a = a.copy()
        m = np.abs(a) < 1e30
        if m.any():
            a[m & (np.abs(a) <= 1e-12 * np.abs(a[m]).max())] = 0.0
        st[f] = a
    return (prm, st), noise


def _run(cls, case, noise, shift=None, eps=0.0, tan=None):
    st = dict(case[1])
    if shift:
        for k, d in shift.items():
            st[k] = st[k] + eps * d
    o = _start(cls, (case[0], st))
    if tan:
        for k, d in tan.items():
            o.set_tan(k, d)
    for m in range(NSTEPS):
        o.step(m, noise[m])
    return o


def test_values_follow_the_plain_oracle(case):
    case_, noise = case
    a = _run(T.Oracle, case_, noise)
    b = _run(T.OracleCS, case_, noise, tan=_direction(1, case_[1]))
    for f in OUT:
        x, y = a.get(f), b.get(f)
        m = np.abs(x) < 1e30
        assert T.rel_err(y, x, m) <= 1e-10, (f, T.rel_err(y, x, m))


def test_tangent_matches_central_differences(case):
    """The tangent against central differences of the plain oracle at a step
    of 1e-5 along the (field-scaled) direction, relative to each field's
    largest slope.  99% of the cells agree to 1e-6 (most to 1e-8; rhop, a
    difference of two values of 1e3, to 2e-7 from the differences' own
    roundoff); the rest are kinks inside the step (a limiter flag, an |x| or
    the min/max of the MPDATA bounds), where a difference quotient mixes the
    slopes either side and the tangent takes one of them; those stay below
    1e-5.  ss_v has one more: the square root of the Richardson
    cutoff (smagorinsky_split_mixing-revised2014.f:112), whose slope is
    unbounded at the cutoff, so a few cells there differ by more."""
    case_, noise = case
    d = _direction(2, case_[1])
    tl = _run(T.OracleCS, case_, noise, tan=d)
    f0 = _run(T.Oracle, case_, noise)
    eps = 1e-5
    fp = _run(T.Oracle, case_, noise, shift=d, eps=eps)
    fm = _run(T.Oracle, case_, noise, shift=d, eps=-eps)
    for f in OUT:
        a0, t = f0.get(f), tl.get_tan(f)
        m = np.abs(a0) < 1e30
        ce = (fp.get(f) - fm.get(f))[m] / (2 * eps)
        scale = np.abs(ce).max()
        e = np.abs(t[m] - ce) / scale
        q99 = np.quantile(e, 0.99)
        far = np.mean(e > 1e-5)
        print("%s: 99%% of cells within %.1e, max %.1e, beyond 1e-5 %.4f" % (f, q99, e.max(), far))
        assert q99 <= 1e-6, (f, q99)
        if f == "ss_v":
            assert far <= 0.01, (f, far)
        else:
            assert e.max() <= 1e-5, (f, e.max())
        assert np.abs(t[m]).max() > 0.0, f


def test_tangent_is_linear(case):
    case_, noise = case
    d1, d2 = _direction(3, case_[1]), _direction(4, case_[1])
    a = _run(T.OracleCS, case_, noise, tan=d1)
    b = _run(T.OracleCS, case_, noise, tan=d2)
    c = _run(T.OracleCS, case_, noise, tan={k: 2.0 * d1[k] - 3.0 * d2[k] for k in d1})
    for f in OUT:
        x = 2.0 * a.get_tan(f) - 3.0 * b.get_tan(f)
        y = c.get_tan(f)
        m = np.abs(a.get(f)) < 1e30
        assert T.rel_err(y, x, m) <= 1e-12, (f, T.rel_err(y, x, m))


def test_plain_build_has_no_tangent():
    prm, before, _, _, _ = load_case(GOLDEN)
    o = T.Oracle(prm)
    with pytest.raises(KeyError):
        o.set_tan("u", before["u"])
