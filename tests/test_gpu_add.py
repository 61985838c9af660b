"""Sums of multiwavelet functions with the node sums on the device (mwg_add,
mwg_tree_add).

Pins:
- on the union grid (prec < 0) the sum is exact linear algebra: point values
  a f(r) + b g(r) and the norm identity |a f + b g|^2 = a^2 |f|^2 + b^2 |g|^2
  + 2 a b <f, g> (norms from every node's wavelet parts, so a wrong branch
  block fails it) to rounding, with one input many levels deeper than the
  other (scaling parts reconstructed along long paths);
- built adaptively (prec >= 0) the sum of two projected Gaussian sums agrees
  with the projection of the joint Gaussian list within a small multiple of
  prec, and an exact cancellation f - f stays at the roots;
- the in-place add equals the out-of-place one.
"""
import numpy as np
import pytest

import mrchem_amd as mw
from mrchem_amd import workloads

pytestmark = pytest.mark.gpu


def _mra(order, max_depth=14, basis="interpolating"):
    return mw.MultiResolutionAnalysis(order=order, basis=basis, max_depth=max_depth, **workloads.WORLD)


def _diff_norm(u, v):
    return np.sqrt(max(u.norm() ** 2 + v.norm() ** 2 - 2.0 * u.dot(v), 0.0))


NARROW = ([1.0], [4000.0], [[0.31, -0.22, 0.13]])
BROAD = ([2.0, -0.7], [0.4, 1.5], [[-0.5, 0.4, 0.0], [0.3, 0.2, -0.6]])


@pytest.mark.parametrize("order,basis", [(5, "interpolating"), (7, "legendre")])
def test_union_grid_sum_is_exact(ctx, order, basis):
    mra = _mra(order, basis=basis)
    f = mw.FunctionTree(mra).project_gaussians(1e-7, *NARROW)
    g = mw.FunctionTree(mra).project_gaussians(1e-4, *BROAD)
    fd, gd = f.nodes()[0].max(), g.nodes()[0].max()
    a, b = 2.0, -0.5
    out = mw.add(-1.0, mw.FunctionTree(mra), [a, b], [f, g], ctx=ctx)
    # union grid: every node of either input, nothing else
    keys = lambda t: {(int(d),) + tuple(int(x) for x in l) for d, l in zip(*t.nodes()[:2])}
    assert keys(out) == keys(f) | keys(g)
    pts = np.random.default_rng(5).uniform(-0.3, 0.3, size=(300, 3)) + np.array(NARROW[2][0])
    pts = np.vstack([pts, np.random.default_rng(6).uniform(-2, 2, size=(100, 3))])
    ref = a * f.evaluate(pts) + b * g.evaluate(pts)
    pv = np.max(np.abs(out.evaluate(pts) - ref)) / np.max(np.abs(ref))
    sq = a * a * f.norm() ** 2 + b * b * g.norm() ** 2 + 2 * a * b * f.dot(g)
    ne = abs(out.norm() ** 2 - sq) / sq
    print("order %d %s: depths %d / %d, %d nodes, point error %.2e, norm identity %.2e" %
          (order, basis, fd, gd, out.size(), pv, ne))
    assert fd - gd >= 5
    assert pv < 1e-12 and ne < 1e-12


def test_adaptive_sum_against_projection(ctx):
    mra = _mra(7)
    prec = 1e-6
    rng = workloads.SplitMix64(4242)
    f1 = workloads.gaussian_sum(rng, 3, 0.5, 40.0, 2.0)
    f2 = workloads.gaussian_sum(rng, 3, 0.5, 40.0, 2.0)
    u = mw.FunctionTree(mra).project_gaussians(prec / 10, *f1)
    v = mw.FunctionTree(mra).project_gaussians(prec / 10, *f2)
    out = mw.add(prec, mw.FunctionTree(mra), [1.0, 1.0], [u, v], abs_prec=True, ctx=ctx)
    joint = tuple(np.concatenate([x, y]) for x, y in zip(f1, f2))
    ref = mw.FunctionTree(mra).project_gaussians(prec / 10, *joint)
    err = _diff_norm(out, ref)
    union = mw.add(-1.0, mw.FunctionTree(mra), [1.0, 1.0], [u, v], ctx=ctx)
    print("adaptive sum: %d nodes (union %d), |out - exact| %.2e (|exact| %.3e)" %
          (out.size(), union.size(), err, ref.norm()))
    assert err < 20 * prec * max(1.0, ref.norm())
    assert out.size() <= union.size() + 8 * union.size() // 4


def test_cancellation_and_relative_precision(ctx):
    mra = _mra(5)
    f = mw.FunctionTree(mra).project_gaussians(1e-5, *BROAD)
    g = mw.FunctionTree(mra).project_gaussians(1e-5, *BROAD)
    zero = mw.add(1e-5, mw.FunctionTree(mra), [1.0, -1.0], [f, g], ctx=ctx)
    n_roots = int(np.prod(workloads.WORLD["boxes"]))
    assert zero.size() == n_roots and zero.norm() < 1e-13
    # relative precision: the same sum scaled by 1000 builds the same grid
    s1 = mw.add(1e-5, mw.FunctionTree(mra), [1.0, 0.3], [f, g], ctx=ctx)
    s2 = mw.add(1e-5, mw.FunctionTree(mra), [1000.0, 300.0], [f, g], ctx=ctx)
    assert s1.size() == s2.size()
    assert abs(s2.norm() / s1.norm() - 1000.0) < 1e-9


def test_in_place_add(ctx):
    mra = _mra(6)
    f = mw.FunctionTree(mra).project_gaussians(1e-6, *NARROW)
    g = mw.FunctionTree(mra).project_gaussians(1e-5, *BROAD)
    ref = mw.add(-1.0, mw.FunctionTree(mra), [1.0, -2.5], [f, g], ctx=ctx)
    f.add(-2.5, g, ctx=ctx)
    assert f.size() == ref.size()
    assert _diff_norm(f, ref) < 1e-12 * ref.norm()
    h = mw.FunctionTree(mra).project_gaussians(1e-5, *BROAD)
    n0 = h.norm()
    h.add(1.0, h, ctx=ctx)
    assert abs(h.norm() - 2 * n0) < 1e-12 * n0


def test_add_argument_errors(ctx):
    mra = _mra(3, max_depth=8)
    other = _mra(4, max_depth=8)
    f = mw.FunctionTree(mra).project_gaussians(1e-3, *BROAD)
    g = mw.FunctionTree(other).project_gaussians(1e-3, *BROAD)
    with pytest.raises(mw.MWGError):
        mw.add(-1.0, f, [1.0], [f], ctx=ctx)
    with pytest.raises(mw.MWGError):
        mw.add(-1.0, mw.FunctionTree(mra), [1.0, 1.0], [f, g], ctx=ctx)
    with pytest.raises(mw.MWGError):
        mw.add(-1.0, mw.FunctionTree(mra), [float("inf")], [f], ctx=ctx)
    with pytest.raises(ValueError):
        mw.add(-1.0, mw.FunctionTree(mra), [1.0, 2.0], [f], ctx=ctx)
