"""Adaptive projection of an analytic function in the interpolating basis
(test helper, numpy).

A node's scaling coefficients are the function's values at its Gauss-Legendre
points times h^{3/2} sqrt(w_i w_j w_k) (h: the node's physical side), so a
node's compressed block [s, d1..d7] is the compression of its 8 children's
point values.  From the roots, a node whose wavelet norm exceeds the split
threshold prec * tnorm * 2^{-(n+1)/2} (WaveletAdaptor's rule, n the node's
scale) is split, up to max_depth; afterwards every branch's block is
recomputed from its children's scaling parts (BottomUp), as MRCPP does after a
projection.  Returns (depth, l, blocks) for FunctionTree.set_nodes."""
import numpy as np

from mw_numpy_ref import compress


def _points(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return (x + 1.0) / 2.0, w / 2.0


def _child_values(f, mra, depth, l, q, sw):
    """Scaling blocks of the 8 children of node (depth, l): [8][z][y][x]."""
    n = mra.kp1
    h = mra.box_size(depth + 1)
    out = np.empty((8, n, n, n))
    for t in range(8):
        c = np.array([2 * l[0] + (t & 1), 2 * l[1] + ((t >> 1) & 1), 2 * l[2] + (t >> 2)], dtype=np.float64)
        xs, ys, zs = (c[0] + q) * h, (c[1] + q) * h, (c[2] + q) * h
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        out[t] = f(X, Y, Z) * sw * h ** 1.5
    return out


def project(mra, f, prec, abs_prec=False, max_depth=None, norm=None):
    """f(X, Y, Z) -> values (numpy broadcasting); norm: ||f|| for the relative
    rule (taken from the result when None: one extra pass)."""
    n = mra.kp1
    F = mra.filter()
    q, w = _points(n)
    sw = np.sqrt(np.einsum("k,j,i->kji", w, w, w))
    max_depth = mra.max_depth if max_depth is None else max_depth
    n3 = n ** 3
    tnorm = 1.0 if abs_prec else (1.0 if norm is None else norm)
    lo, hi = [c for c in mra.corner], [c + b for c, b in zip(mra.corner, mra.boxes)]
    roots = [(0, (x, y, z)) for z in range(lo[2], hi[2]) for y in range(lo[1], hi[1]) for x in range(lo[0], hi[0])]
    nodes = {}
    work = roots
    while work:
        nxt = []
        for d, l in work:
            blk = compress(_child_values(f, mra, d, l, q, sw), F, n)
            nodes[(d, l)] = blk
            wn = np.sqrt(np.sum(blk[1:] ** 2))
            scale = mra.min_scale + d
            if d < max_depth and wn > max(2 * 2.220446049250313e-16, prec * tnorm * 2.0 ** (-0.5 * (scale + 1))):
                nxt += [(d + 1, (2 * l[0] + (t & 1), 2 * l[1] + ((t >> 1) & 1), 2 * l[2] + (t >> 2))) for t in range(8)]
        work = nxt
    # BottomUp: every branch from its children's scaling parts
    for d in sorted({k[0] for k in nodes}, reverse=True):
        for (dd, l) in [k for k in nodes if k[0] == d]:
            kids = [(d + 1, (2 * l[0] + (t & 1), 2 * l[1] + ((t >> 1) & 1), 2 * l[2] + (t >> 2))) for t in range(8)]
            if kids[0] in nodes:
                ch = np.stack([nodes[k][0] for k in kids])
                nodes[(d, l)] = compress(ch, F, n)
    keys = sorted(nodes, key=lambda k: k[0])
    depth = np.array([k[0] for k in keys], dtype=np.int32)
    ls = np.array([k[1] for k in keys], dtype=np.int64)
    blocks = np.stack([nodes[k].reshape(-1) for k in keys])
    return depth, ls, blocks
