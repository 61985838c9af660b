"""Packed trees (wire and file format, mwg_tree_pack / unpack / save / load):
no GPU needed.

Pins: the documented layout (header fields at their offsets, the padded
depth array, translations, blocks in tree order) read back with struct and
numpy independently of the library; bit-identical round trips through memory
and through a file; refusal of another MRA, a truncated or resized buffer and
a bad magic; zero-initialised nodes and round trips of an 80 MB tree
(the huge-page allocation path)."""
import struct

import numpy as np
import pytest

import mrchem_amd as mw
from mrchem_amd import workloads


def _tree(order=4, basis="interpolating"):
    mra = mw.MultiResolutionAnalysis(order=order, basis=basis, max_depth=10, **workloads.WORLD)
    f = mw.FunctionTree(mra).project_gaussians(1e-5, [1.0, 0.5], [20.0, 3.0], [[0.2, 0.1, -0.3], [-0.4, 0.0, 0.2]])
    return mra, f


def test_layout_matches_documentation():
    mra, f = _tree()
    b = f.pack()
    d, l, c = f.nodes()
    n, blk = len(d), c.shape[1]
    magic, version, kp1, basis, n0 = struct.unpack_from("<8s4i", b, 0)
    corner = struct.unpack_from("<3i", b, 24)
    boxes = struct.unpack_from("<3i", b, 36)
    sfac, n_nodes, sq = struct.unpack_from("<dqd", b, 48)
    assert magic == b"MWGTREE1" and version == 1 and kp1 == 5 and n0 == workloads.WORLD["min_scale"]
    assert corner == tuple(workloads.WORLD["corner"]) and boxes == tuple(workloads.WORLD["boxes"])
    assert sfac == workloads.WORLD["scaling_factor"] and n_nodes == n
    assert sq == pytest.approx(f.square_norm(), rel=1e-14)
    off = 72
    dep = np.frombuffer(b, dtype="<i4", count=n, offset=off)
    off += (4 * n + 7) // 8 * 8
    trans = np.frombuffer(b, dtype="<i8", count=3 * n, offset=off).reshape(n, 3)
    off += 24 * n
    blocks = np.frombuffer(b, dtype="<f8", count=n * blk, offset=off).reshape(n, blk)
    assert off + 8 * n * blk == len(b)
    assert np.array_equal(dep, d) and np.array_equal(trans, l) and np.array_equal(blocks, c)
    # parents before children
    seen = set()
    for k in zip(dep, *trans.T):
        k = tuple(int(x) for x in k)
        if k[0] > 0:
            assert (k[0] - 1,) + tuple(x >> 1 for x in k[1:]) in seen
        seen.add(k)


@pytest.mark.parametrize("basis", ["interpolating", "legendre"])
def test_round_trips_are_bit_identical(tmp_path, basis):
    mra, f = _tree(basis=basis)
    g = mw.FunctionTree(mra).unpack(f.pack())
    for x, y in zip(f.nodes(), g.nodes()):
        assert np.array_equal(x, y)
    assert g.square_norm() == f.square_norm()
    path = tmp_path / "phi_idx_0_re.tree"
    f.save(path)
    h = mw.FunctionTree(mra).load(path)
    for x, y in zip(f.nodes(), h.nodes()):
        assert np.array_equal(x, y)
    assert path.read_bytes() == f.pack()


def test_refuses_bad_input():
    mra, f = _tree()
    b = f.pack()
    other = mw.MultiResolutionAnalysis(order=5, basis="interpolating", max_depth=10, **workloads.WORLD)
    with pytest.raises(mw.MWGError, match="different MRA"):
        mw.FunctionTree(other).unpack(b)
    with pytest.raises(mw.MWGError, match="size"):
        mw.FunctionTree(mra).unpack(b[:-8])
    with pytest.raises(mw.MWGError, match="size"):
        mw.FunctionTree(mra).unpack(b + b"\0" * 8)
    with pytest.raises(mw.MWGError, match="truncated"):
        mw.FunctionTree(mra).unpack(b[:40])
    with pytest.raises(mw.MWGError, match="magic"):
        mw.FunctionTree(mra).unpack(b"X" + b[1:])
    with pytest.raises(mw.MWGError):
        mw.FunctionTree(mra).load("/nonexistent/dir/tree.bin")


def test_large_tree_zero_init_and_round_trip():
    """Trees whose coefficients take 4 MiB or more are allocated in 2 MiB
    aligned, huge-page-advised blocks: nodes created without coefficients
    still read as zeros there, and an 80 MB tree packs, unpacks and sets
    back bit-identically."""
    cfg = workloads.config("c2")
    mra = mw.MultiResolutionAnalysis(order=cfg["order"], basis="interpolating", max_depth=16, **workloads.WORLD)
    c, a, R = cfg["densities"][0]
    f = mw.FunctionTree(mra).project_gaussians(cfg["prec"], c, a, R)
    d, l, blocks = f.nodes()
    assert blocks.nbytes > 64 << 20
    empty = mw.FunctionTree(mra).set_nodes(d, l)
    assert not np.any(empty.nodes()[2])
    g = mw.FunctionTree(mra).unpack(f.pack())
    for x, y in zip(f.nodes(), g.nodes()):
        assert np.array_equal(x, y)
    h = mw.FunctionTree(mra).set_nodes(d, l, blocks)
    assert np.array_equal(h.nodes()[2], blocks) and h.square_norm() == f.square_norm()
