"""A run set up from files (case.open_case: case file + stratification
table) starts from the oracle's initial state: INITIALIZE_PROFILES, DENSITY,
INITIALIZE and SPONGE_MAKER (hotcross80-revised2014.f:210-255)."""
import os
import sys

import numpy as np
import pytest

import hc_testlib as T
import gpu_common as G
from test_case import odd_table

sys.path.insert(0, os.path.join(T.ROOT, "plumemodel-hotcross_amd"))
import case  # noqa: E402

pytestmark = pytest.mark.gpu
RHO_ULPS = 2 * float(np.spacing(1030.0))


def test_open_case_matches_oracle_initial_state(tmp_path):
    print("open_case 16x12x10", flush=True)
    prm = T.reference_params(16, 12, 10)
    cpath, tpath = str(tmp_path / "run.case"), str(tmp_path / "strat.inc")
    case.write_case(G.to_hc(prm), cpath, {"title": "golden grid"})
    odd_table(tpath)
    ctx, extras = case.open_case(cpath, tpath)
    try:
        assert extras == {"title": "golden grid"}
        o = T.Oracle(prm)
        o.init_profiles(*case.read_stratification(tpath))
        for c in ("density", "initialize", "sponge_maker"):
            o.call(c)
        for f in ("u", "v", "w", "um", "vm", "wm", "up", "vp", "wp", "t", "s", "c", "p"):
            a, b = ctx.download(f), o.get(f)
            assert np.array_equal(a, b), f
        assert G.compare_abs(ctx.download("rhop"), o.get("rhop")) <= RHO_ULPS
    finally:
        ctx.close()
