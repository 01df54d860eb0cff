"""The autotuner's device hooks and the compiled table: recorded contraction
shapes report the plan the library uses, which for a shape in the table is
the table's, and a forced plan whose split-K partials do not fit the call's
workspace is refused before anything is queued."""
import os
import re

import pytest

import mpsamd as M
from mpsamd import core
from mpsamd.trajectory import SpinEnsemble

pytestmark = pytest.mark.gpu

INC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "matrixproductstates.jl_amd", "csrc",
                   "gemm_tuned.inc")


def _table():
    out = {}
    for ln in open(INC).read().splitlines():
        m = re.match(r"\s*\{(\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+), (\d+)\},\s*//\s*([^:]*):", ln)
        if m:
            v = [int(x) for x in m.groups()[:8]]
            out[tuple(v[:6])] = (v[6], v[7], m.group(9))
    return out


def test_recorded_plans_are_the_tables():
    table = _table()
    label = "spin D=16 w=4 d=2 nb=96 fp32"
    assert any(t[2] == label for t in table.values())
    print("case recorded plans", label, flush=True)
    ens = SpinEnsemble(60, 16, 1, list(range(96)), dt=0.01, qr_rank_tol=1e-13)
    ens.ctx.set_precision("fp32")
    ens.step()
    ens.ctx.set_graphs(False)
    core.gemm_record(ens.ctx, True)
    ens.step()
    core.gemm_record(ens.ctx, False)
    shapes = core.gemm_recorded(ens.ctx)
    assert len(shapes) >= 5
    hit = 0
    for s in shapes:
        key = (s["M"], s["N"], s["K"], s["batch"], s["nb_class"], s["cmp"])
        # the FP32 form: operator terms in FP32 (cmp 1), overlap terms in FP64 (cmp 0)
        assert s["tm"] in (32, 64) and s["splits"] >= 1 and s["nb"] == 96 and s["cmp"] in (0, 1)
        if key in table:
            assert (s["tm"], s["splits"]) == table[key][:2], key
            hit += 1
    assert hit >= 1 and any(s["cmp"] == 1 for s in shapes)
    # timing a recorded shape, and a split-K depth beyond the workspace
    assert core.gemm_time(ens.ctx, 0, shapes[0]["tm"], shapes[0]["splits"], 3) > 0.0
    from mpsamd.autotune import _effective_splits
    over = [r for r, s in enumerate(shapes)
            if 0 < s["ws_bytes"] < _effective_splits(s["K"], 64) * s["batch"] * s["M"] * s["N"] * 16]
    fits = [r for r, s in enumerate(shapes)
            if _effective_splits(s["K"], 2) * s["batch"] * s["M"] * s["N"] * 16 <= s["ws_bytes"]]
    print("shapes", len(shapes), "beyond the workspace at depth 64:", len(over), flush=True)
    for r in over[:2]:
        with pytest.raises(M.MPSError) as e:
            core.gemm_time(ens.ctx, r, 32, 64, 1)
        assert e.value.code == -1
    for r in fits[:2]:
        assert core.gemm_time(ens.ctx, r, 64, 2, 2) > 0.0
    ens.close()


def test_forced_plan_beyond_the_workspace_is_refused():
    # an 8-site d = 9 chain at D = 128 has the 1792 x 1152 products whose
    # partials at depth 64 (8 chunks of K = 128) need far more than the call's workspace
    from mpsamd.autotune import _effective_splits, rydberg_single
    print("case forced plan beyond the workspace", flush=True)
    ctx, step, close = rydberg_single(N=8, D=128, prec="fp64")
    try:
        step()
        ctx.synchronize()
        ctx.set_graphs(False)
        core.gemm_record(ctx, True)
        step()
        ctx.synchronize()
        core.gemm_record(ctx, False)
        shapes = core.gemm_recorded(ctx)
        over = [r for r, s in enumerate(shapes)
                if 0 < s["ws_bytes"] < _effective_splits(s["K"], 64) * s["batch"] * s["M"] * s["N"] * 16]
        assert over
        with pytest.raises(M.MPSError) as e:
            core.gemm_time(ctx, over[0], 64, 64, 1)
        assert e.value.code == -1
        # the recorded call's own plan still runs
        s = shapes[over[0]]
        assert core.gemm_time(ctx, over[0], s["tm"], s["splits"], 2) > 0.0
    finally:
        close()
