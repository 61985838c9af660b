"""GPU apply vs the FP64 oracle on identical frozen grids and thresholds.

The HIP path (C ABI, libmwgreen.so) computes the output grid adaptively; the
oracle recomputes every node on that grid with the same per-node thresholds,
so the screened term sets coincide and the FP64 results must agree to
rounding (relative 1e-12 of the largest coefficient).
"""
import math

import numpy as np
import pytest

import mrchem_amd as mw
import oracle
from oracle import closed_forms as cf

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


def _mra(order, basis="interpolating", max_depth=16):
    return mw.MultiResolutionAnalysis(order=order, basis=basis, min_scale=-4, scaling_factor=1.25,
                                      max_depth=max_depth)


def _parity(mra, op, rho, V):
    od, ol, oc = V.nodes()
    ref, st = oracle.apply_fixed(mra, op.tables(), rho.nodes(), od, ol, V.thresholds())
    err = np.max(np.abs(ref - oc)) / np.max(np.abs(ref))
    return err, st


@pytest.mark.parametrize("order,basis", [(5, "interpolating"), (7, "interpolating"), (4, "legendre")])
def test_poisson_single_gaussian(ctx, order, basis):
    mra = _mra(order, basis)
    prec = 1e-4 if order < 7 else 1e-5
    P = mw.PoissonOperator(mra, prec)
    a, r0 = 10.0, np.array([0.31, -0.17, 0.52])
    rho = mw.FunctionTree(mra).project_gaussians(prec, [1.0], [a], r0[None, :])
    V = mw.FunctionTree(mra)
    mw.apply(prec / rho.norm(), V, P, rho, ctx=ctx)
    err, st = _parity(mra, P, rho, V)
    print("order %d %s: nodes %d apps %.3e err %.2e" % (order, basis, V.size(), st["term_apps"], err))
    assert err < TOL_FP64
    E = rho.dot(V)
    E_ex = math.sqrt(2 * a / math.pi)
    # the adaptive output grid (split rule ||d|| > prec ||V|| 2^{-(n+1)/2})
    # leaves an energy error of 5-15 prec at these orders
    assert abs(E - E_ex) / E_ex < 30 * prec


def test_helmholtz_gaussian_sum(ctx):
    mra = _mra(6)
    prec, mu = 1e-4, 1.2
    rng = np.random.default_rng(7)
    n = 4
    coef = rng.normal(size=n)
    exps = np.exp(rng.uniform(np.log(0.5), np.log(20.0), size=n))
    cen = rng.uniform(-1.5, 1.5, size=(n, 3))
    H = mw.HelmholtzOperator(mra, mu, prec)
    rho = mw.FunctionTree(mra).project_gaussians(prec, coef, exps, cen)
    V = mw.FunctionTree(mra)
    mw.apply(prec, V, H, rho, abs_prec=True, ctx=ctx)
    err, st = _parity(mra, H, rho, V)
    assert err < TOL_FP64
    E = rho.dot(V)
    E_ex = cf.pair_energy(coef, exps, cen, mu=mu)
    assert abs(E - E_ex) < 30 * prec * max(1.0, abs(E_ex))


@pytest.mark.parametrize("order", [7, 9])
def test_mixed_precision_term_set_and_error(ctx, order):
    """Mixed mode applies exactly the FP64 term set (screening identical, the
    FP32 class is the complement of the FP64 class) and stays within a small
    fraction of prec of the FP64 result on the same frozen grid."""
    mra = _mra(order)
    prec = 1e-6
    P = mw.PoissonOperator(mra, prec)
    rng = np.random.default_rng(11)
    n = 4
    coef = rng.normal(size=n)
    exps = np.exp(rng.uniform(np.log(1.0), np.log(60.0), size=n))
    cen = rng.uniform(-1.0, 1.0, size=(n, 3))
    rho = mw.FunctionTree(mra).project_gaussians(prec, coef, exps, cen)
    V = mw.FunctionTree(mra)
    mw.apply(prec / rho.norm(), V, P, rho, ctx=ctx)
    thr = V.thresholds()
    od, ol, _ = V.nodes()
    V2 = mw.FunctionTree(mra).set_nodes(od, ol)
    p64 = mw.ApplyPlan(V, P, rho, thr, mode="fp64", ctx=ctx)
    pmx = mw.ApplyPlan(V2, P, rho, thr, mode="mixed", ctx=ctx)
    r64 = p64.execute()
    rmx = pmx.execute()
    assert r64["term_apps"] == rmx["term_apps"]
    assert r64["term_apps"] == p64.info()["host_term_apps"]
    assert 0 < rmx["lp_term_apps"] < rmx["term_apps"]
    ns64 = p64.download_ns()
    nsmx = pmx.download_ns()
    err = np.max(np.abs(nsmx - ns64)) / np.max(np.abs(ns64))
    print("order %d: %d of %d term-apps on FP32, max coef err vs FP64 %.2e" %
          (order, rmx["lp_term_apps"], rmx["term_apps"], err))
    # requirement: within prec of the FP64 result, relative to the largest
    # coefficient (observed ~1e-7 at tau = 65536)
    assert err < prec
    # and the FP64 plan reproduces the oracle on this grid
    ref, st = oracle.apply_fixed(mra, P.tables(), rho.nodes(), od, ol, thr, want_ns=False)
    p64.finish()
    c64 = V.nodes()[2]
    assert np.max(np.abs(ref - c64)) / np.max(np.abs(ref)) < TOL_FP64
    assert st["term_apps"] == r64["term_apps"]


@pytest.mark.parametrize("order,prec,basis", [(2, 1e-4, "interpolating"), (3, 1e-4, "interpolating"),
                                              (4, 1e-5, "interpolating"), (5, 1e-5, "interpolating"),
                                              (7, 1e-6, "interpolating"), (8, 1e-6, "interpolating"),
                                              (9, 1e-6, "interpolating"), (10, 1e-5, "interpolating"),
                                              (11, 1e-5, "interpolating"), (5, 1e-5, "legendre"),
                                              (9, 1e-6, "legendre"), (11, 1e-5, "legendre")])
def test_split16_term_set_and_error(ctx, order, prec, basis):
    """FP16 hi/lo split products: same term set as FP64 and within prec of
    the FP64 result on the identical grid (N <= 8, N = 9..10 and N = 11..12
    take different operand layouts; N = 9 and 11 have partly empty tails), in
    both bases; and the FP64 plan equal to the oracle on that grid at every
    N (the FP64 kernel's tile split differs by N)."""
    mra = _mra(order, basis)
    P = mw.PoissonOperator(mra, prec)
    rng = np.random.default_rng(23)
    n = 4
    coef = rng.normal(size=n)
    exps = np.exp(rng.uniform(np.log(1.0), np.log(60.0), size=n))
    cen = rng.uniform(-1.0, 1.0, size=(n, 3))
    rho = mw.FunctionTree(mra).project_gaussians(prec, coef, exps, cen)
    V = mw.FunctionTree(mra)
    mw.apply(prec / rho.norm(), V, P, rho, ctx=ctx)
    thr = V.thresholds()
    od, ol, _ = V.nodes()
    V2 = mw.FunctionTree(mra).set_nodes(od, ol)
    p64 = mw.ApplyPlan(V, P, rho, thr, mode="fp64", ctx=ctx)
    psp = mw.ApplyPlan(V2, P, rho, thr, mode="split16", ctx=ctx)
    r64 = p64.execute()
    rsp = psp.execute()
    assert r64["term_apps"] == rsp["term_apps"]
    assert 0 < rsp["lp_term_apps"] < rsp["term_apps"]
    ns64 = p64.download_ns()
    nssp = psp.download_ns()
    err = np.max(np.abs(nssp - ns64)) / np.max(np.abs(ns64))
    print("order %d %s: %d of %d term-apps as FP16 split products, max coef err vs FP64 %.2e" %
          (order, basis, rsp["lp_term_apps"], rsp["term_apps"], err))
    assert err < prec
    p64.finish()
    psp.finish()
    E64, Esp = rho.dot(V), rho.dot(V2)
    assert abs(Esp - E64) < prec * max(1.0, abs(E64))
    # the FP64 plan (pair jobs; each N has its own tile split) against the
    # oracle on this grid
    ref, st = oracle.apply_fixed(mra, P.tables(), rho.nodes(), od, ol, thr, want_ns=False)
    c64 = V.nodes()[2]
    assert np.max(np.abs(ref - c64)) / np.max(np.abs(ref)) < TOL_FP64
    assert st["term_apps"] == r64["term_apps"]


@pytest.mark.parametrize("order", [1, 9, 10, 11])
def test_post_transforms_on_wide_levels(ctx, order):
    """A frozen grid whose root box around the density is complete to depth
    4 (the other root boxes are leaves): its levels hold 1, 8, 64 and 512
    branches.  Levels of up to 128 branches (64 at k = 11) run the post
    transforms as one workgroup per (branch, output block), the 512-branch
    level as one workgroup per branch (k <= 9) or per (branch, first-pass
    half) (k = 10, 11), so two TopDown/BottomUp kernels meet in one tree; the
    post-transformed FP64 result equals the oracle's on the same grid and
    thresholds."""
    mra = _mra(order)
    prec = 1e-4
    P = mw.PoissonOperator(mra, prec)
    cen = np.array([0.3, -0.2, 0.1])
    rho = mw.FunctionTree(mra).project_gaussians(prec, [1.0], [10.0], cen[None, :])
    d0, l0, _ = rho.nodes()
    roots = l0[d0 == 0]
    # the root box holding the centre (root boxes are 2^4 x 1.25 = 20 wide
    # here); any root would do, this one carries the largest values
    want = np.floor(cen / 20.0).astype(np.int64)
    hit = np.where(np.all(roots == want, axis=1))[0]
    box = roots[hit[0]] if len(hit) else roots[0]
    depth, keys = [0] * len(roots), [np.asarray(roots)]
    for d in range(1, 5):
        g = np.stack(np.meshgrid(*[np.arange(2 ** d)] * 3, indexing="ij"), -1).reshape(-1, 3)
        keys.append(box[None, :] * 2 ** d + g)
        depth += [d] * len(g)
    od, ol = np.array(depth, dtype=np.int32), np.concatenate(keys).astype(np.int64)
    assert len(od) == len(roots) + 8 + 64 + 512 + 4096
    V = mw.FunctionTree(mra).set_nodes(od, ol)
    t = 1e-3 * rho.norm()
    plan = mw.ApplyPlan(V, P, rho, np.full(len(od), t), mode="fp64", ctx=ctx)
    r = plan.execute()
    plan.finish()
    vd, vl, vc = V.nodes()
    ref, st = oracle.apply_fixed(mra, P.tables(), rho.nodes(), vd, vl, np.full(len(vd), t))
    assert st["term_apps"] == r["term_apps"] > 0
    err = np.max(np.abs(ref - vc)) / np.max(np.abs(ref))
    print("order %d: %d nodes, %d term-apps, post-transformed FP64 vs oracle %.2e" %
          (order, len(vd), r["term_apps"], err))
    assert err < TOL_FP64
