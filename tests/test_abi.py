"""The C ABI boundary: symbols, argument validation and host-side trees.

No GPU work happens here; these run on the CPU-only CI host.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import mrchem_amd as mw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "mwgreen.h")).read()
    return sorted(set(re.findall(r"\b(mwg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    syms = _declared_symbols()
    assert len(syms) >= 25
    lib = ctypes.CDLL(mw.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_library_has_gfx950_code_object():
    data = open(mw.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_mra_argument_errors():
    with pytest.raises(mw.MWGError):
        mw.MultiResolutionAnalysis(order=0)
    with pytest.raises(mw.MWGError):
        mw.MultiResolutionAnalysis(order=5, scaling_factor=-1.0)
    with pytest.raises(mw.MWGError):
        mw.MultiResolutionAnalysis(order=5, max_depth=40)


def test_operator_argument_errors():
    mra = mw.MultiResolutionAnalysis(order=4, max_depth=10)
    with pytest.raises(mw.MWGError):
        mw.PoissonOperator(mra, 0.0)
    with pytest.raises(mw.MWGError):
        mw.HelmholtzOperator(mra, -1.0, 1e-4)


def test_tree_set_roundtrip_and_validation():
    mra = mw.MultiResolutionAnalysis(order=3, min_scale=-4, scaling_factor=1.25, max_depth=10)
    f = mw.FunctionTree(mra).project_gaussians(1e-4, [1.0], [5.0], [[0.1, 0.2, 0.3]])
    d, l, c = f.nodes()
    perm = np.random.default_rng(1).permutation(len(d))
    g = mw.FunctionTree(mra).set_nodes(d[perm], l[perm], c[perm])
    assert abs(g.dot(f) - f.square_norm()) < 1e-12 * f.square_norm()
    # dropping one child of a family is an error, not a crash
    child = np.where(d == d.max())[0][0]
    keep = np.ones(len(d), bool)
    keep[child] = False
    with pytest.raises(mw.MWGError):
        mw.FunctionTree(mra).set_nodes(d[keep], l[keep], c[keep])
    with pytest.raises(mw.MWGError):
        mw.FunctionTree(mra).set_nodes(d, l + 10 ** 6, c)


def test_apply_rejects_bad_arguments_without_device():
    mra = mw.MultiResolutionAnalysis(order=3, max_depth=8)
    P = mw.PoissonOperator(mra, 1e-3)
    f = mw.FunctionTree(mra).project_gaussians(1e-3, [1.0], [2.0], [[0, 0, 0]])
    out = mw.FunctionTree(mra)
    null_ctx = mw.Context.__new__(mw.Context)
    null_ctx._h = None
    with pytest.raises(mw.MWGError):
        mw.apply(1e-3, out, P, f, ctx=null_ctx)
    other = mw.MultiResolutionAnalysis(order=3, max_depth=8)
    with pytest.raises(ValueError):
        mw.apply_fixed(out, P, f, np.zeros(out.size() + 1), ctx=null_ctx)


def test_operator_tables_layout():
    mra = mw.MultiResolutionAnalysis(order=4, min_scale=-4, scaling_factor=1.25, max_depth=12)
    P = mw.PoissonOperator(mra, 1e-4)
    t = P.tables()
    kp1 = 5
    # every stored block's norm matches its Frobenius norm
    for i in range(0, t["M"] * t["ndepth"], 37):
        L = t["L"][i]
        if L < 0:
            continue
        nb = 4 * (2 * L + 1)
        blk = t["blocks"][t["boff"][i]:t["boff"][i] + nb * kp1 * kp1].reshape(nb, kp1 * kp1)
        nr = t["norms"][t["noff"][i]:t["noff"][i] + nb]
        assert np.allclose(np.linalg.norm(blk, axis=1), nr, rtol=1e-12, atol=0)


def test_cxx_caller_links():
    """The C++ caller (tests/cxx/mwg_caller.cpp, built by build()) includes
    include/mwgreen.hpp and links libmwgreen.so; --link-only resolves the
    ABI entry points it calls without touching a device."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cxx", "mwg_caller")
    assert os.path.exists(exe), "tests/cxx/mwg_caller is not built; run __graft_entry__.build()"
    r = subprocess.run([exe, "--link-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mwg_apply_batch" in r.stdout
