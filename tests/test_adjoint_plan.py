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
