"""Handle lifetimes, the per-context error and the whole-apply plan step.

- A live plan holds device pointers into its operators' tables and writes
  into its output trees: destroying the operator, an output tree or the
  context while the plan lives is an argument error that leaves the handle
  usable; after the plan is destroyed the same calls succeed.
- mwg_plan_download_ns checks the caller's capacity; mwg_plan_finish refuses
  an output tree whose grid changed since the plan was made.
- Failures on a context are also kept per context (mwg_context_last_error).
- mwg_plan_apply (calculator + TopDown/BottomUp + norms on the device, in
  place) gives the same finished tree as mwg_plan_execute + mwg_plan_finish,
  and repeated applies give the same result (each step re-zeroes the output).
"""
import ctypes

import numpy as np
import pytest

import mrchem_amd as mw

pytestmark = pytest.mark.gpu


def _setup(ctx, order=5, prec=1e-4):
    mra = mw.MultiResolutionAnalysis(order=order, basis="interpolating", min_scale=-4, scaling_factor=1.25,
                                     max_depth=12)
    P = mw.PoissonOperator(mra, prec)
    rho = mw.FunctionTree(mra).project_gaussians(prec, [1.0, 0.5], [8.0, 3.0], [[0.1, 0.2, -0.1], [-0.4, 0.0, 0.3]])
    V = mw.FunctionTree(mra)
    mw.apply(prec, V, P, rho, abs_prec=True, ctx=ctx)
    return mra, P, rho, V


def test_destroy_under_live_plan_is_an_argument_error(ctx):
    mra, P, rho, V = _setup(ctx)
    d, l, _ = V.nodes()
    out = mw.FunctionTree(mra).set_nodes(d, l)
    plan = mw.ApplyPlan(out, P, rho, V.thresholds(), mode="fp64", ctx=ctx)
    assert mw.lib.mwg_operator_destroy(P._h) == -1
    assert "plan" in mw.lib.mwg_last_error().decode()
    assert mw.lib.mwg_tree_destroy(out._h) == -1
    assert mw.lib.mwg_context_destroy(ctx._h) == -1
    # all three still work
    plan.execute()
    plan.finish()
    assert out.norm() > 0.0 and P.size() > 0
    V2 = mw.FunctionTree(mra)
    mw.apply(1e-4, V2, P, rho, abs_prec=True, ctx=ctx)
    assert V2.size() == V.size()
    # once the plan is gone, the operator and the tree can be destroyed
    del plan
    tmp = mw.PoissonOperator(mra, 1e-3)
    t2 = mw.FunctionTree(mra).set_nodes(d, l)
    p2 = mw.ApplyPlan(t2, tmp, rho, V.thresholds(), mode="fp64", ctx=ctx)
    del p2
    h, tmp._h = tmp._h, None
    assert mw.lib.mwg_operator_destroy(h) == 0
    h, t2._h = t2._h, None
    assert mw.lib.mwg_tree_destroy(h) == 0


def test_download_capacity_and_changed_grid(ctx):
    mra, P, rho, V = _setup(ctx)
    d, l, _ = V.nodes()
    out = mw.FunctionTree(mra).set_nodes(d, l)
    plan = mw.ApplyPlan(out, P, rho, V.thresholds(), mode="fp64", ctx=ctx)
    plan.execute()
    n = plan.info()["out_nodes"]
    small = np.zeros(n * mra.block - 1)
    rc = mw.lib.mwg_plan_download_ns(plan._h, small.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), small.size)
    assert rc == -1 and "holds" in mw.lib.mwg_last_error().decode()
    assert "holds" in ctx.last_error()
    ns = plan.download_ns()
    assert ns.shape == (n, mra.block) and np.any(ns)
    # the same number of nodes on another grid: the 8 children of one branch
    # whose children are all leaves move under a leaf elsewhere
    d2, l2 = _moved_children(d, l)
    assert len(d2) == len(d)
    out.set_nodes(d2, l2)
    with pytest.raises(mw.MWGError, match="grid changed"):
        plan.finish()


def _moved_children(d, l):
    keys = {(int(a), *map(int, b)) for a, b in zip(d, l)}
    parent = [(int(a) - 1, *(int(x) >> 1 for x in b)) if a > 0 else None for a, b in zip(d, l)]
    branch = {p for p in parent if p is not None}
    leaf_children = {}
    for i, p in enumerate(parent):
        if p is not None:
            leaf_children.setdefault(p, []).append(i)
    p1 = next(p for p, ch in leaf_children.items() if all((int(d[i]), *map(int, l[i])) not in branch for i in ch))
    drop = set(leaf_children[p1])
    dropped = {(int(d[i]), *map(int, l[i])) for i in drop}
    p2 = next(k for k in sorted(keys) if k not in branch and k != p1 and k not in dropped)
    keep = [i for i in range(len(d)) if i not in drop]
    nd = list(d[keep]) + [p2[0] + 1] * 8
    nl = [list(l[i]) for i in keep] + [[2 * p2[1] + (c & 1), 2 * p2[2] + ((c >> 1) & 1), 2 * p2[3] + (c >> 2)]
                                       for c in range(8)]
    return np.array(nd, dtype=np.int32), np.array(nl, dtype=np.int64)


def test_context_last_error_is_per_context(ctx):
    other = mw.Context(0)
    with pytest.raises(mw.MWGError):
        other.set_precision("fp64", 0.5)
    assert "ratio" in other.last_error()
    assert "ratio" not in ctx.last_error()


@pytest.mark.parametrize("mode", ["fp64", "split16"])
def test_plan_apply_equals_execute_and_finish(ctx, mode):
    mra, P, rho, V = _setup(ctx, order=7, prec=1e-5)
    d, l, _ = V.nodes()
    thr = V.thresholds()
    a = mw.FunctionTree(mra).set_nodes(d, l)
    b = mw.FunctionTree(mra).set_nodes(d, l)
    pa = mw.ApplyPlan(a, P, rho, thr, mode=mode, ctx=ctx)
    pb = mw.ApplyPlan(b, P, rho, thr, mode=mode, ctx=ctx)
    r1 = pa.apply()
    r2 = pa.apply()
    assert r1["term_apps"] == r2["term_apps"] and r1["post_ms"] > 0.0
    with pytest.raises(mw.MWGError, match="post transforms"):
        pa.download_ns()
    pa.finish()
    pb.execute()
    pb.finish()
    ca, cb = a.nodes()[2], b.nodes()[2]
    err = np.max(np.abs(ca - cb)) / np.max(np.abs(cb))
    print("%s: apply %.3f ms (post %.3f ms) vs execute+finish: max rel diff %.2e" %
          (mode, r2["ms"], r2["post_ms"], err))
    # the same arithmetic; only the order of the FP64 atomic sums differs
    assert err < 1e-12
    assert abs(a.square_norm() - b.square_norm()) <= 1e-12 * b.square_norm()
    if mode == "fp64":
        # and the FP64 result is the adaptive apply's (same grid and thresholds)
        assert np.max(np.abs(ca - V.nodes()[2])) / np.max(np.abs(ca)) < 1e-13
