"""Products of multiwavelet functions on the device (mwg_multiply).

Pins:
- the product of two projected Gaussian sums against the projection of the
  exact product (Gaussian product theorem, workloads.pair_density): norm of
  the difference and point values within a small multiple of prec;
- the node math against a numpy restatement: for an output node, both
  inputs' scaling blocks at its 8 children (reconstructed from the inputs'
  nodes, wavelet parts zero below an input's finest node), the interpolating
  point product c s^a s^b h^{-3/2} / sqrt(w_i w_j w_k), and the compression
  of the 8 children.
"""
import numpy as np
import pytest

import mrchem_amd as mw
from mrchem_amd import workloads
from mw_numpy_ref import compress, reconstruct, scaling_at

pytestmark = pytest.mark.gpu


def _mra(order, max_depth=14):
    return mw.MultiResolutionAnalysis(order=order, basis="interpolating", max_depth=max_depth, **workloads.WORLD)


def _orbitals(seed, n, n_gauss, lo, hi, box):
    rng = workloads.SplitMix64(seed)
    return [workloads.gaussian_sum(rng, n_gauss, lo, hi, box) for _ in range(n)]


def _diff_norm(u, v):
    return np.sqrt(max(u.norm() ** 2 + v.norm() ** 2 - 2.0 * u.dot(v), 0.0))


@pytest.mark.parametrize("order,prec", [(5, 1e-5), (7, 1e-6)])
def test_product_of_gaussian_sums(ctx, order, prec):
    mra = _mra(order)
    fi, fj = _orbitals(9090 + order, 2, 3, 0.5, 20.0, 2.0)
    a = mw.FunctionTree(mra).project_gaussians(prec / 10, *fi)
    b = mw.FunctionTree(mra).project_gaussians(prec / 10, *fj)
    out = mw.multiply(prec, mw.FunctionTree(mra), a, b, abs_prec=True, ctx=ctx)
    ref = mw.FunctionTree(mra).project_gaussians(prec / 10, *workloads.pair_density(fi, fj))
    err = _diff_norm(out, ref)
    pts = np.random.default_rng(3).uniform(-2.5, 2.5, size=(200, 3))
    pv = np.max(np.abs(out.evaluate(pts) - ref.evaluate(pts)))
    scale = np.max(np.abs(ref.evaluate(pts)))
    print("order %d: %d nodes (inputs %d, %d), |out - exact| = %.2e (|exact| %.3e), max point error %.2e of %.2e" %
          (order, out.size(), a.size(), b.size(), err, ref.norm(), pv, scale))
    assert err < 20 * prec * max(1.0, ref.norm())
    assert pv < 100 * prec * max(1.0, scale)


def test_node_math_against_numpy(ctx):
    order, prec = 3, 1e-4
    n = order + 1
    mra = _mra(order, max_depth=10)
    fi = ([1.0], [8.0], [[0.3, -0.2, 0.1]])
    fj = ([0.7, -0.4], [2.0, 30.0], [[-0.5, 0.4, 0.0], [0.2, 0.2, -0.3]])
    a = mw.FunctionTree(mra).project_gaussians(prec, *fi)
    b = mw.FunctionTree(mra).project_gaussians(prec, *fj)
    coef = 1.7
    out = mw.multiply(prec, mw.FunctionTree(mra), a, b, coef=coef, max_iter=0, abs_prec=True, ctx=ctx)
    F = mra.filter()
    x, w = np.polynomial.legendre.leggauss(n)
    w = w / 2.0
    wsc = 1.0 / np.sqrt(np.einsum("k,j,i->kji", w, w, w))
    an, bn = a.nodes(), b.nodes()
    od, ol, oc = out.nodes()
    leaves = [i for i in range(len(od)) if not np.any((od == od[i] + 1) & np.all((ol >> 1) == ol[i], axis=1))]
    # errors relative to the product's largest coefficient: far from both
    # Gaussians the reconstructed values are tiny differences of larger
    # ancestor coefficients, and only their absolute error is meaningful
    checked = deps = 0
    errs, scale = [], 0.0
    for i in leaves[:: max(1, len(leaves) // 64)]:
        depth, l = int(od[i]), ol[i]
        sa, sb = scaling_at(an, depth, l, F, n), scaling_at(bn, depth, l, F, n)
        ca, cb = reconstruct(sa, F, n), reconstruct(sb, F, n)
        h = mra.box_size(depth + 1)
        p = coef * ca * cb * h ** -1.5 * wsc[None]
        ref = compress(p, F, n).ravel()
        errs.append(np.max(np.abs(oc[i] - ref)))
        scale = max(scale, np.max(np.abs(ref)))
        checked += 1
        deps += int(not np.any((an[0] == depth) & np.all(an[1] == l, axis=1)))
    err = max(errs) / scale
    print("checked %d output leaves (%d below an input's finest node), max error %.2e of the largest coefficient" %
          (checked, deps, err))
    assert checked >= 16 and deps >= 1
    assert err < 1e-13


def test_device_against_oracle(ctx):
    """Every leaf of an adaptive device product against the oracle's product
    restatement at the same nodes (oracle.multiply, pinned on the CPU by the
    identity f x 1 = f and the numpy restatement)."""
    import oracle
    mra = _mra(5)
    fi, fj = _orbitals(4711, 2, 3, 0.5, 20.0, 2.0)
    a = mw.FunctionTree(mra).project_gaussians(1e-6, *fi)
    b = mw.FunctionTree(mra).project_gaussians(1e-6, *fj)
    out = mw.multiply(1e-5, mw.FunctionTree(mra), a, b, coef=0.7, abs_prec=True, ctx=ctx)
    od, ol, oc = out.nodes()
    leaves = np.array([not np.any((od == od[i] + 1) & np.all((ol >> 1) == ol[i], axis=1)) for i in range(len(od))])
    ref = oracle.multiply(mra, a.nodes(), b.nodes(), od[leaves], ol[leaves], coef=0.7)
    err = np.max(np.abs(oc[leaves] - ref)) / np.max(np.abs(ref))
    print("%d leaves of %d nodes: device vs oracle %.2e" % (leaves.sum(), len(od), err))
    assert leaves.sum() > 100 and err < 1e-13


def test_multiply_argument_errors(ctx):
    mra = _mra(3, max_depth=8)
    f = mw.FunctionTree(mra).project_gaussians(1e-3, [1.0], [2.0], [[0.0, 0.0, 0.0]])
    with pytest.raises(mw.MWGError):
        mw.multiply(1e-3, f, f, f, ctx=ctx)
    leg = mw.MultiResolutionAnalysis(order=3, basis="legendre", max_depth=8, **workloads.WORLD)
    g = mw.FunctionTree(leg).project_gaussians(1e-3, [1.0], [2.0], [[0.0, 0.0, 0.0]])
    with pytest.raises(mw.MWGError, match="interpolating"):
        mw.multiply(1e-3, mw.FunctionTree(leg), g, g, ctx=ctx)
