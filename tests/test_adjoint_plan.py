"""The checkpoint-interval planner of the adjoint window (adjoint.plan_segment)
on stand-in contexts: it returns the interval with the smallest footprint
and that footprint, and refuses a window that does not fit the budget.  The
256^3 FP32 figure the design notes quote (1800 steps: a checkpoint every 30
steps, 134 GB) follows from the window's storage rule, (ceil(n/seg) + 1 + seg)
states of 14 fields plus (seg - 1) steps of 2 x 8 tracer fields, with the
padded 256^3 field of 288 x 259 x 259 FP32 values."""
import math
import os
import sys

import pytest

import hc_testlib as T

sys.path.insert(0, os.path.join(T.ROOT, "plumemodel-hotcross_amd"))
import adjoint as AD  # noqa: E402


class Footprint:
    def __init__(self, fn):
        self.fn = fn
        self.calls = 0

    def adj_window_bytes(self, nsteps, seg):
        self.calls += 1
        return self.fn(nsteps, seg)


def window_rule(field_bytes):
    return lambda n, s: (math.ceil(n / s) + 1 + s) * 14 * field_bytes + (s - 1) * 16 * field_bytes


def test_planner_takes_the_smallest_footprint():
    ctx = Footprint(lambda n, s: (s - 7) ** 2 + 100)
    seg, need = AD.plan_segment(ctx, 20, 1e9)
    assert (seg, need) == (7, 100)


def test_planner_refuses_an_oversized_window():
    ctx = Footprint(lambda n, s: 1000 + s)
    with pytest.raises(ValueError, match="over the"):
        AD.plan_segment(ctx, 10, 500)


def test_headline_window_figure():
    fb = 288 * 259 * 259 * 4
    seg, need = AD.plan_segment(Footprint(window_rule(fb)), 1800, 200e9)
    assert seg == 30
    assert abs(need / 1e9 - 134.3) < 0.1
    # the balance point of 14 n / s against 30 s
    assert abs(seg - math.sqrt(14 * 1800 / 30)) < 2


# ---------------------------------------------------------------------------
# binomial checkpointing: the schedule of hc_binomial_schedule (host only)

import hotcross as H  # noqa: E402
from functools import lru_cache  # noqa: E402


@lru_cache(None)
def _advances(l, s):
    """Fewest advances to reverse l steps from a held start state with s free
    slots, by brute force over every split (a turn costs no advance)."""
    if l == 1:
        return 0
    if s == 0:
        return l * (l - 1) // 2
    return min(m + _advances(l - m, s - 1) + _advances(m, s) for m in range(1, l))


def _run_schedule(nsteps, nsnap):
    """Executes the schedule on a model of the window: slot contents and the
    live state as step numbers.  Checks every action's precondition and
    returns the turned steps in order and the advances."""
    acts, nfwd = H.binomial_schedule(nsteps, nsnap)
    slots = {}
    live = 0
    turned, adv = [], 0
    for kind, step, slot in acts:
        if kind == 0:  # advance
            assert live == step, (live, step)
            live = step + 1
            adv += 1
        elif kind == 1:  # snapshot of the live state
            assert live == step and 0 <= slot <= nsnap
            slots[slot] = step
        elif kind == 2:  # restore
            assert slots.get(slot) == step
            live = step
        else:  # turn step n from X_n in slot; the live state must be X_n
            assert slots.get(slot) == step and live == step, (slot, step, live)
            turned.append(step)
            live = None  # the buffers hold X_n+1 under X_n's host state: restore first
    assert adv + len(turned) == nfwd
    return turned, adv


@pytest.mark.parametrize("nsteps,nsnap", [(1, 1), (2, 1), (7, 1), (8, 2), (8, 3), (13, 3), (20, 4), (25, 6), (30, 30),
                                          (31, 2), (64, 5)])
def test_binomial_schedule_reverses_every_step_once(nsteps, nsnap):
    turned, adv = _run_schedule(nsteps, nsnap)
    assert turned == list(range(nsteps - 1, -1, -1))
    # the fewest advances for nsnap slots (slot 0 holds the start)
    assert adv == _advances(nsteps, nsnap - 1)


@pytest.mark.parametrize("nsteps,nsnap", [(1800, 35), (1800, 124), (400, 10), (100, 1)])
def test_binomial_schedule_at_window_sizes(nsteps, nsnap):
    """At the forward + adjoint configuration's length the advances follow
    the closed form t l - beta(s+2, t-1) (s = nsnap - 1, t the least with
    beta(s+1, t) >= l); slot 0 and nsnap - 1 further slots, plus the
    turns' temporary, are all the schedule uses."""
    turned, adv = _run_schedule(nsteps, nsnap)
    assert turned == list(range(nsteps - 1, -1, -1))
    s = nsnap - 1
    if s == 0:
        want = nsteps * (nsteps - 1) // 2
    else:
        t = 1
        while math.comb(s + 1 + t, t) < nsteps:
            t += 1
        want = t * nsteps - math.comb(s + 2 + t - 1, t - 1)
    assert adv == want
    acts, _ = H.binomial_schedule(nsteps, nsnap)
    assert acts[:, 2].max() <= nsnap


class TwoSchemes:
    """Stand-in context: the fixed window's and the binomial window's storage
    rules of the library at 256^3 FP32 (field 288 x 259 x 259 x 4 bytes)."""
    fb = 288 * 259 * 259 * 4

    def adj_window_bytes(self, n, s):
        return (math.ceil(n / s) + 1 + s) * 14 * self.fb + (s - 1) * 16 * self.fb

    def adj_window_binomial_bytes(self, n, nsnap):
        return (nsnap + 2) * 14 * self.fb + 16 * self.fb


def test_window_planner_falls_back_to_binomial():
    ctx = TwoSchemes()
    kind, seg, need = AD.plan_window(ctx, 1800, 200e9)
    assert kind == "fixed" and seg == 30 and need <= 200e9
    kind, nsnap, need = AD.plan_window(ctx, 1800, 60e9)
    assert kind == "binomial" and need <= 60e9
    assert ctx.adj_window_binomial_bytes(1800, nsnap + 1) > 60e9
    with pytest.raises(ValueError):
        AD.plan_window(ctx, 1800, 2e9)
