"""numpy restatement of the two-scale node transforms (test helper).

Blocks are [8][z][y][x]; component t has bit d set for a wavelet along d
(x: bit 0, y: bit 1, z: bit 2).  F is the MRA's 2(k+1) x 2(k+1) filter
[H0 H1; G0 G1] (mra.filter())."""
import numpy as np


# numpy restatement of the node transforms (component t: bit d set = wavelet
# along d; blocks indexed [z][y][x])
def pair_apply(blk, F, n, dim, transpose):
    out = blk.copy()
    M = F.T if transpose else F
    axis = 3 - dim  # block axes: (t, z, y, x)
    for t in range(8):
        if (t >> dim) & 1:
            continue
        lo, hi = blk[t], blk[t | (1 << dim)]
        v = np.concatenate([lo, hi], axis=axis - 1)
        u = np.tensordot(M, np.moveaxis(v, axis - 1, 0), axes=(1, 0))
        u = np.moveaxis(u, 0, axis - 1)
        sl = [slice(None)] * 3
        sl[axis - 1] = slice(0, n)
        out[t] = u[tuple(sl)]
        sl[axis - 1] = slice(n, 2 * n)
        out[t | (1 << dim)] = u[tuple(sl)]
    return out


def reconstruct(blk, F, n):
    for dim in (2, 1, 0):
        blk = pair_apply(blk, F, n, dim, True)
    return blk


def compress(blk, F, n):
    for dim in (0, 1, 2):
        blk = pair_apply(blk, F, n, dim, False)
    return blk


def child_scaling(blk, F, n, t):
    """Scaling block of child t of a full node block (reconstruct, pick)."""
    return reconstruct(blk, F, n)[t]


def scaling_at(tree_nodes, depth, l, F, n):
    """Scaling block of (depth, l) from the tree's deepest node on its path."""
    d_arr, l_arr, c_arr = tree_nodes
    for d in range(depth, -1, -1):
        lp = np.array(l) >> (depth - d)
        hit = np.where((d_arr == d) & np.all(l_arr == lp, axis=1))[0]
        if len(hit) == 0:
            continue
        blk = c_arr[hit[0]].reshape(8, n, n, n)
        for lv in range(d + 1, depth + 1):
            bits = (np.array(l) >> (depth - lv)) & 1
            t = int(bits[0] + 2 * bits[1] + 4 * bits[2])
            s = child_scaling(blk, F, n, t)
            blk = np.zeros_like(blk)
            blk[0] = s
        return blk
    raise AssertionError("no source node")
