"""The oracle's product restatement (oracle.multiply), pinned on the CPU.

- Identity: f times the constant function 1 (a tree of roots only) gives
  f's own blocks at every node of f, to rounding.  This fixes the point-value
  scaling (h^{3/2} sqrt(w_i w_j w_k)), the pairing of reconstruction and
  compression, and the descent through many levels below an input's finest
  node (the constant tree has no node below the roots).
- A constant c times g is c g, and the product is symmetric in its inputs.
- Products of two Gaussian sums at the leaves of a grid against the numpy
  restatement of the node transforms (tests/mw_numpy_ref.py), including
  leaves below one input's finest node.
"""
import numpy as np

import mrchem_amd as mw
import oracle
from mrchem_amd import workloads
from mw_numpy_ref import compress, reconstruct, scaling_at


def _mra(order):
    return mw.MultiResolutionAnalysis(order=order, basis="interpolating", max_depth=10, **workloads.WORLD)


def _constant(mra, c):
    """The constant c as a tree of roots: s_ijk = c h^{3/2} sqrt(w_i w_j w_k)."""
    n = mra.kp1
    _, w = np.polynomial.legendre.leggauss(n)
    w = w / 2.0
    s = c * mra.box_size(0) ** 1.5 * np.sqrt(np.einsum("k,j,i->kji", w, w, w)).ravel()
    boxes, corner = workloads.WORLD["boxes"], workloads.WORLD["corner"]
    keys = [(corner[0] + x, corner[1] + y, corner[2] + z)
            for z in range(boxes[2]) for y in range(boxes[1]) for x in range(boxes[0])]
    blocks = np.zeros((len(keys), mra.block))
    blocks[:, :n ** 3] = s
    return mw.FunctionTree(mra).set_nodes(np.zeros(len(keys), dtype=np.int32), np.array(keys), blocks)


GA = ([1.0, -0.3], [12.0, 2.5], [[0.2, -0.1, 0.05], [-0.4, 0.3, 0.1]])
GB = ([0.8], [30.0], [[0.1, 0.1, -0.2]])


def test_times_one_is_identity():
    for order in (3, 6):
        mra = _mra(order)
        f = mw.FunctionTree(mra).project_gaussians(1e-6, *GA)
        one = _constant(mra, 1.0)
        d, l, c = f.nodes()
        out = oracle.multiply(mra, f.nodes(), one.nodes(), d, l)
        err = np.max(np.abs(out - c)) / np.max(np.abs(c))
        print("order %d: f x 1 vs f over %d nodes (depth <= %d): %.2e" % (order, len(d), d.max(), err))
        assert d.max() >= 4 and err < 1e-13


def test_constant_scaling_and_symmetry():
    mra = _mra(4)
    g = mw.FunctionTree(mra).project_gaussians(1e-5, *GB)
    three = _constant(mra, 3.0)
    d, l, c = g.nodes()
    a = oracle.multiply(mra, three.nodes(), g.nodes(), d, l, coef=0.5)
    b = oracle.multiply(mra, g.nodes(), three.nodes(), d, l, coef=0.5)
    scale = np.max(np.abs(c))
    assert np.max(np.abs(a - 1.5 * c)) < 1e-13 * scale
    assert np.max(np.abs(a - b)) < 1e-15 * scale


def test_gaussian_products_against_numpy():
    mra = _mra(3)
    n = mra.kp1
    a = mw.FunctionTree(mra).project_gaussians(1e-4, *GA)
    b = mw.FunctionTree(mra).project_gaussians(1e-4, *GB)
    an, bn = a.nodes(), b.nodes()
    # output nodes: b's leaves (deeper than a around b's centre) and a's
    bd, bl, _ = bn
    leaves = [i for i in range(len(bd)) if not np.any((bd == bd[i] + 1) & np.all((bl >> 1) == bl[i], axis=1))]
    od, ol = bd[leaves], bl[leaves]
    out = oracle.multiply(mra, an, bn, od, ol, coef=1.3)
    F = mra.filter()
    _, w = np.polynomial.legendre.leggauss(n)
    w = w / 2.0
    wsc = 1.0 / np.sqrt(np.einsum("k,j,i->kji", w, w, w))
    errs, scale, deeper = [], 0.0, 0
    for q in range(0, len(od), max(1, len(od) // 40)):
        depth, l = int(od[q]), ol[q]
        ca = reconstruct(scaling_at(an, depth, l, F, n), F, n)
        cb = reconstruct(scaling_at(bn, depth, l, F, n), F, n)
        ref = compress(1.3 * ca * cb * mra.box_size(depth + 1) ** -1.5 * wsc[None], F, n).ravel()
        errs.append(np.max(np.abs(out[q] - ref)))
        scale = max(scale, np.max(np.abs(ref)))
        deeper += int(not np.any((an[0] == depth) & np.all(an[1] == l, axis=1)))
    print("%d leaves checked (%d below a's finest node): max error %.2e of %.2e" % (len(errs), deeper, max(errs), scale))
    assert deeper >= 1 and max(errs) < 1e-14 * scale
