"""FunctionTree::crop on the host tree (mwg_tree_crop): no GPU needed.

Pins: cropping removes only children of branch nodes whose wavelet norm is
below the threshold, so the function changes by at most about the sum of the
dropped wavelet parts (checked in norm and at points); it is idempotent; a
tiny precision keeps every node; the tree stays complete (every kept node's
parent is kept, children in groups of 8)."""
import numpy as np

import mrchem_amd as mw
from mrchem_amd import workloads

GAUSS = ([1.0, -0.4], [30.0, 2.0], [[0.1, -0.2, 0.3], [-0.4, 0.1, 0.0]])


def _tree(prec=1e-7, order=5):
    mra = mw.MultiResolutionAnalysis(order=order, basis="interpolating", max_depth=12, **workloads.WORLD)
    return mra, mw.FunctionTree(mra).project_gaussians(prec, *GAUSS)


def _keys(t):
    d, l, _ = t.nodes()
    return {(int(a),) + tuple(int(x) for x in b) for a, b in zip(d, l)}


def _complete(t):
    ks = _keys(t)
    for k in ks:
        if k[0] == 0:
            continue
        par = (k[0] - 1,) + tuple(x >> 1 for x in k[1:])
        assert par in ks
        sib = [(k[0],) + tuple(2 * p + ((c >> i) & 1) for i, p in enumerate(par[1:])) for c in range(8)]
        assert all(s in ks for s in sib)


def test_crop_shrinks_within_tolerance():
    mra, f = _tree()
    ref = mw.FunctionTree(mra).project_gaussians(1e-7, *GAUSS)
    n0 = f.size()
    f.crop(1e-3)
    assert f.size() < n0 and _keys(f) <= _keys(ref)
    _complete(f)
    diff = np.sqrt(max(f.norm() ** 2 + ref.norm() ** 2 - 2 * f.dot(ref), 0.0))
    pts = np.random.default_rng(1).uniform(-1, 1, size=(200, 3))
    pv = np.max(np.abs(f.evaluate(pts) - ref.evaluate(pts)))
    print("crop 1e-3: %d -> %d nodes, |diff| %.2e, max point diff %.2e" % (n0, f.size(), diff, pv))
    assert diff < 1e-2 and diff > 0
    n1 = f.size()
    f.crop(1e-3)
    assert f.size() == n1


def test_crop_tiny_precision_keeps_every_node():
    _, f = _tree(prec=1e-4)
    ks = _keys(f)
    f.crop(1e-15)
    assert _keys(f) == ks


def test_crop_relative_matches_scaled_absolute():
    _, f = _tree(prec=1e-6)
    _, g = _tree(prec=1e-6)
    f.crop(1e-3, abs_prec=False)
    g.crop(1e-3 * g.norm(), abs_prec=True)
    assert _keys(f) == _keys(g)
