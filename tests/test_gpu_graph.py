"""Replayed hipGraphs of the frozen-grid step (mwg_plan_use_graph).

The step (output and counters cleared by kernels, the persistent calculator
launch or launches, and for mwg_plan_apply the post transforms and norms) is
captured after one eager run and replayed.  Each replay must give the eager
result: the same term count and the same blocks (to 1e-13 of the largest:
the FP64 sums of a node's segments are atomic, so their order may differ),
with other eager work and synchronisations between the replays.  The
mixed mode's capture forks the FP64 class onto the second stream and joins
it back.
"""
import numpy as np
import pytest

import mrchem_amd as mw

pytestmark = pytest.mark.gpu


def _grid(ctx, order, prec):
    mra = mw.MultiResolutionAnalysis(order=order, basis="interpolating", min_scale=-4, scaling_factor=1.25,
                                     max_depth=14)
    H = mw.HelmholtzOperator(mra, 1.1, prec)
    rho = mw.FunctionTree(mra).project_gaussians(prec, [1.0, -0.6, 0.3], [9.0, 2.5, 30.0],
                                                  [[0.1, 0.2, -0.1], [-0.4, 0.0, 0.3], [0.3, -0.2, 0.1]])
    V = mw.FunctionTree(mra)
    mw.apply(prec, V, H, rho, abs_prec=True, ctx=ctx)
    return mra, H, rho, V


@pytest.mark.parametrize("mode", ["fp64", "mixed", "split16"])
def test_graph_replays_equal_eager(ctx, mode):
    mra, H, rho, V = _grid(ctx, 7, 1e-6)
    d, l, _ = V.nodes()
    thr = V.thresholds()
    plan = mw.ApplyPlan(mw.FunctionTree(mra).set_nodes(d, l), H, rho, thr, mode=mode, ctx=ctx)
    other = mw.ApplyPlan(mw.FunctionTree(mra).set_nodes(d, l), H, rho, thr, mode="fp64", ctx=ctx)
    r0 = plan.execute()
    ref = plan.download_ns()
    scale = np.max(np.abs(ref))
    plan.use_graph(True)
    for rep in range(4):
        other.execute()  # eager work on the same streams between replays
        r = plan.execute()
        ns = plan.download_ns()
        err = np.max(np.abs(ns - ref)) / scale
        print("%s replay %d: %.3f ms (eager %.3f ms), %d term-apps, max diff %.2e" %
              (mode, rep, r["ms"], r0["ms"], r["term_apps"], err))
        assert r["term_apps"] == r0["term_apps"] and r["lp_term_apps"] == r0["lp_term_apps"]
        assert err < 1e-13
    # the whole apply (calculator + post transforms + norms) as one graph
    a = mw.FunctionTree(mra).set_nodes(d, l)
    pa = mw.ApplyPlan(a, H, rho, thr, mode=mode, ctx=ctx)
    pa.apply()
    pa.finish()
    eager = a.nodes()[2].copy()
    eager_sq = a.square_norm()
    pa.use_graph(True)
    for rep in range(3):
        other.apply()
        r = pa.apply()
        pa.finish()
        c = a.nodes()[2]
        err = np.max(np.abs(c - eager)) / np.max(np.abs(eager))
        print("%s apply replay %d: %.3f ms, max diff %.2e" % (mode, rep, r["ms"], err))
        assert err < 1e-13
        assert abs(a.square_norm() - eager_sq) <= 1e-12 * eager_sq
