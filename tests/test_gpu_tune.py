"""The autotuner of frozen-grid plans (mwg_plan_tune).

For a split16 and a mixed plan: the tuner's pick keeps the largest output
difference to an FP64 run of the same plan within the requested bound, the
term set is unchanged, the reported time of the pick is the fastest it saw
among the settings within the bound (never slower than the starting setting
when that one is within it), and a new plan of the same
(k + 1, terms, band width, mode) starts from the cached pick.
"""
import numpy as np
import pytest

import mrchem_amd as mw

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["split16", "mixed"])
def test_tune_keeps_error_bound_and_caches(mode):
    ctx = mw.Context(0)  # its own cache of tuned knobs
    mra = mw.MultiResolutionAnalysis(order=7, basis="interpolating", min_scale=-4, scaling_factor=1.25,
                                     max_depth=14)
    prec = 1e-6
    H = mw.HelmholtzOperator(mra, 1.3, prec)
    rho = mw.FunctionTree(mra).project_gaussians(prec, [1.0, -0.7, 0.4], [12.0, 3.0, 40.0],
                                                  [[0.1, 0.2, -0.1], [-0.4, 0.0, 0.3], [0.2, -0.3, 0.1]])
    V = mw.FunctionTree(mra)
    mw.apply(prec, V, H, rho, abs_prec=True, ctx=ctx)
    d, l, _ = V.nodes()
    thr = V.thresholds()
    p64 = mw.ApplyPlan(mw.FunctionTree(mra).set_nodes(d, l), H, rho, thr, mode="fp64", ctx=ctx)
    r64 = p64.execute()
    ns64 = p64.download_ns()
    scale = np.max(np.abs(ns64))
    plan = mw.ApplyPlan(mw.FunctionTree(mra).set_nodes(d, l), H, rho, thr, mode=mode, ctx=ctx)
    bound = 0.1 * prec
    rep = plan.tune(bound, reps=3)
    print("%s tune: %s" % (mode, rep))
    assert rep["rel_err"] <= bound and rep["tried"] >= 3
    if rep["start_rel_err"] <= bound:
        assert rep["ms"] <= rep["start_ms"] and rep["speedup"] >= 1.0
    r = plan.execute()
    assert r["term_apps"] == r64["term_apps"]
    err = np.max(np.abs(plan.download_ns() - ns64)) / scale
    assert err <= bound * 1.01, err
    # a new plan with the same key starts from the cached pick
    again = mw.ApplyPlan(mw.FunctionTree(mra).set_nodes(d, l), H, rho, thr, mode=mode, ctx=ctx)
    inf = again.info()
    assert (inf["tau"], inf["single_ratio"], inf["order"]) == (rep["tau"], rep["single_ratio"], rep["order"])
    rep2 = again.tune(bound, reps=1)
    print("%s second tune (from the cached pick): %s" % (mode, rep2))
    assert rep2["rel_err"] <= bound
    r2 = again.execute()
    assert r2["term_apps"] == r64["term_apps"]
    assert np.max(np.abs(again.download_ns() - ns64)) / scale <= bound * 1.01
