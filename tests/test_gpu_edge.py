"""Edge cases of the device apply.

- A zero input function (zero blocks on a refined grid): the output is the
  zero function on the roots, since no node has a wavelet part to refine
  on, in every arithmetic mode.
- The lowest order (k = 1, two basis functions per direction): FP64 parity
  with the oracle on the GPU's grid and thresholds, and the FP16-split plan
  on that grid within prec of it.
- An order the device kernels are not built for (k + 1 = 13): a clear
  argument error from the apply, the frozen-grid plan and the device
  operator build, raised before any device work, and the context stays
  usable afterwards.
"""
import numpy as np
import pytest

import mrchem_amd as mw
import oracle

pytestmark = pytest.mark.gpu


def _mra(order, max_depth=12):
    return mw.MultiResolutionAnalysis(order=order, basis="interpolating", min_scale=-4, scaling_factor=1.25,
                                      max_depth=max_depth)


@pytest.mark.parametrize("mode", ["fp64", "mixed", "split16"])
def test_zero_input_gives_zero_on_the_roots(ctx, mode):
    mra = _mra(5)
    P = mw.PoissonOperator(mra, 1e-4)
    g = mw.FunctionTree(mra).project_gaussians(1e-4, [1.0], [20.0], [[0.1, -0.2, 0.05]])
    d, l, c = g.nodes()
    assert d.max() >= 2
    zero = mw.FunctionTree(mra).set_nodes(d, l, np.zeros_like(c))
    assert zero.norm() == 0.0
    V = mw.FunctionTree(mra)
    try:
        ctx.set_precision(mode)
        mw.apply(1e-4, V, P, zero, abs_prec=True, ctx=ctx)
    finally:
        ctx.set_precision("fp64")
    od, _, oc = V.nodes()
    assert od.max() == 0 and V.size() == np.count_nonzero(d == 0)
    assert not np.any(oc) and V.norm() == 0.0


def test_lowest_order_parity(ctx):
    mra = _mra(1)
    prec = 1e-3
    P = mw.PoissonOperator(mra, prec)
    rho = mw.FunctionTree(mra).project_gaussians(prec, [1.0, -0.5], [3.0, 1.0], [[0.2, 0.1, -0.3], [-0.4, 0.3, 0.2]])
    V = mw.FunctionTree(mra)
    mw.apply(prec, V, P, rho, abs_prec=True, ctx=ctx)
    od, ol, oc = V.nodes()
    thr = V.thresholds()
    ref, st = oracle.apply_fixed(mra, P.tables(), rho.nodes(), od, ol, thr)
    err = np.max(np.abs(ref - oc)) / np.max(np.abs(ref))
    plan = mw.ApplyPlan(mw.FunctionTree(mra).set_nodes(od, ol), P, rho, thr, mode="split16", ctx=ctx)
    r = plan.execute()
    p64 = mw.ApplyPlan(mw.FunctionTree(mra).set_nodes(od, ol), P, rho, thr, mode="fp64", ctx=ctx)
    p64.execute()
    ns64 = p64.download_ns()
    err16 = np.max(np.abs(plan.download_ns() - ns64)) / np.max(np.abs(ns64))
    print("k = 1: %d nodes (depth <= %d), %d term-apps, oracle err %.2e, split16 err %.2e" %
          (V.size(), od.max(), st["term_apps"], err, err16))
    assert od.max() >= 2
    assert err < 1e-12
    assert r["term_apps"] == st["term_apps"] and err16 < prec


def test_unsupported_order_is_an_argument_error(ctx):
    mra = _mra(12, max_depth=6)
    P = mw.PoissonOperator(mra, 1e-3)
    rho = mw.FunctionTree(mra).project_gaussians(1e-3, [1.0], [2.0], [[0.0, 0.0, 0.0]])
    with pytest.raises(mw.MWGError, match=r"k \+ 1 <= 12"):
        mw.apply(1e-3, mw.FunctionTree(mra), P, rho, abs_prec=True, ctx=ctx)
    d, l, _ = rho.nodes()
    with pytest.raises(mw.MWGError, match=r"k \+ 1 <= 12"):
        mw.ApplyPlan(mw.FunctionTree(mra).set_nodes(d, l), P, rho, np.full(len(d), 1e-3), mode="fp64", ctx=ctx)
    with pytest.raises(mw.MWGError, match="unsupported basis size"):
        mw.build_operators(mra, [0.0, 1.0], 1e-3, ctx=ctx)
    # the context still applies at a supported order
    mra5 = _mra(5)
    P5 = mw.PoissonOperator(mra5, 1e-4)
    rho5 = mw.FunctionTree(mra5).project_gaussians(1e-4, [1.0], [5.0], [[0.0, 0.1, 0.0]])
    V5 = mw.FunctionTree(mra5)
    mw.apply(1e-4, V5, P5, rho5, abs_prec=True, ctx=ctx)
    assert V5.size() > 0 and V5.norm() > 0.0


def test_large_download_through_staging_matches_oracle(ctx):
    """Outputs of 32 MiB and more reach the host through two pinned staging
    buffers in 64 MiB chunks.  The Helmholtz workload "c2" (mu 1.2, a sum of
    10 Gaussians, k = 9, prec 1e-6; 2,688 output nodes, 172 MB of
    non-standard blocks: two full chunks and a partial one) downloaded that
    way equals the oracle's raw non-standard blocks on a sample of nodes
    taken from every chunk, including both sides of each chunk boundary and
    the last node."""
    from mrchem_amd import workloads
    cfg = workloads.config("c2")
    mra = mw.MultiResolutionAnalysis(order=cfg["order"], basis="interpolating", max_depth=16, **workloads.WORLD)
    H = mw.HelmholtzOperator(mra, cfg["mu"][0], cfg["prec"])
    c, a, R = cfg["densities"][0]
    f = mw.FunctionTree(mra).project_gaussians(cfg["prec"], c, a, R)
    V = mw.FunctionTree(mra)
    mw.apply(cfg["prec"], V, H, f, abs_prec=True, ctx=ctx)
    od, ol, _ = V.nodes()
    thr = V.thresholds()
    plan = mw.ApplyPlan(mw.FunctionTree(mra).set_nodes(od, ol), H, f, thr, mode="fp64", ctx=ctx)
    plan.execute()
    ns = plan.download_ns()
    node_bytes = 8 * mra.block
    assert len(od) * node_bytes > 2 * (64 << 20)
    pick = set(range(0, len(od), 97)) | {len(od) - 1}
    for b in range(1, len(od) * node_bytes // (64 << 20) + 1):
        q = b * (64 << 20) // node_bytes  # the node that straddles chunk boundary b
        pick |= {q - 1, q, min(q + 1, len(od) - 1)}
    pick = np.array(sorted(pick))
    _, _, ref = oracle.apply_fixed(mra, H.tables(), f.nodes(), od[pick], ol[pick], thr[pick], want_ns=True, post=False)
    err = np.max(np.abs(ns[pick] - ref)) / np.max(np.abs(ref))
    print("%d nodes (%.0f MB), %d sampled: max rel err %.2e" % (len(od), len(od) * node_bytes / 1e6, len(pick), err))
    assert err < 1e-12
