"""Batched applies (mwg_apply_batch, mwg_plan_create_batch) on the GPU.

A batch must give each apply exactly what the single apply gives: the same
adaptive grid and thresholds (each apply's refinement depends only on its own
norms) and the same coefficients, and the FP64 oracle on each apply's frozen
grid must agree to rounding.  The exchange-style batch (pair densities of
Gaussian orbitals, Poisson) and the Helmholtz batch (one mu per orbital) are
checked against their closed-form energies.
"""
import numpy as np
import pytest

import mrchem_amd as mw
import oracle
from mrchem_amd import workloads
from oracle import closed_forms as cf

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


def _mra(order, max_depth=16):
    return mw.MultiResolutionAnalysis(order=order, basis="interpolating", max_depth=max_depth, **workloads.WORLD)


def _orbitals(seed, n, n_gauss, lo, hi, box):
    rng = workloads.SplitMix64(seed)
    return [workloads.gaussian_sum(rng, n_gauss, lo, hi, box) for _ in range(n)]


def test_batch_matches_single_applies(ctx):
    """Poisson pair densities and Helmholtz applies with distinct mu in one
    batch: identical grids, thresholds and coefficients to one-by-one applies."""
    mra = _mra(5)
    prec = 1e-4
    P = mw.PoissonOperator(mra, prec)
    H1 = mw.HelmholtzOperator(mra, 0.8, prec)
    H2 = mw.HelmholtzOperator(mra, 1.9, prec)
    orbs = _orbitals(4242, 3, 3, 0.5, 20.0, 2.0)
    inps = [mw.FunctionTree(mra).project_gaussians(prec, *workloads.pair_density(orbs[0], orbs[1])),
            mw.FunctionTree(mra).project_gaussians(prec, *orbs[1]),
            mw.FunctionTree(mra).project_gaussians(prec, *orbs[2])]
    ops = [P, H1, H2]
    entries = [dict(prec=prec, out=mw.FunctionTree(mra), oper=o, inp=f, abs_prec=True) for o, f in zip(ops, inps)]
    mw.apply_batch(entries, ctx=ctx)
    st = ctx.last_stats()
    for e in entries:
        single = mw.FunctionTree(mra)
        mw.apply(prec, single, e["oper"], e["inp"], abs_prec=True, ctx=ctx)
        db, lb, cb = e["out"].nodes()
        ds, ls, cs = single.nodes()
        assert np.array_equal(db, ds) and np.array_equal(lb, ls)
        assert np.allclose(e["out"].thresholds(), single.thresholds(), rtol=1e-12, atol=0.0)
        err = np.max(np.abs(cb - cs)) / np.max(np.abs(cs))
        assert err < TOL_FP64, err
    print("batch of %d: %d launches, %d output nodes" % (len(entries), st["launches"], st["out_nodes"]))


def test_batch_plan_fp64_and_mixed_against_oracle(ctx):
    """Batched frozen plan: FP64 equals the oracle apply by apply; mixed applies
    the same term set and stays within prec of FP64."""
    mra = _mra(7)
    prec = 1e-5
    P = mw.PoissonOperator(mra, prec)
    orbs = _orbitals(777, 3, 4, 0.5, 30.0, 2.5)
    dens = [workloads.pair_density(orbs[i], orbs[j]) for i in range(3) for j in range(i, 3)][:4]
    inps = [mw.FunctionTree(mra).project_gaussians(prec, *d) for d in dens]
    outs = [mw.FunctionTree(mra) for _ in inps]
    mw.apply_batch([dict(prec=prec, out=o, oper=P, inp=f, abs_prec=True) for o, f in zip(outs, inps)], ctx=ctx)
    thr = [o.thresholds() for o in outs]
    grids = [o.nodes()[:2] for o in outs]
    outs_mx = [mw.FunctionTree(mra).set_nodes(*g) for g in grids]
    p64 = mw.ApplyPlan(outs, [P] * len(outs), inps, thr, mode="fp64", ctx=ctx)
    pmx = mw.ApplyPlan(outs_mx, [P] * len(outs), inps, thr, mode="mixed", ctx=ctx)
    assert p64.info()["applies"] == len(outs)
    r64, rmx = p64.execute(), pmx.execute()
    assert r64["term_apps"] == rmx["term_apps"] == p64.info()["host_term_apps"]
    ns64, nsmx = p64.download_ns(), pmx.download_ns()
    err_mx = np.max(np.abs(nsmx - ns64)) / np.max(np.abs(ns64))
    assert err_mx < prec
    p64.finish()
    total = 0
    for o, f, (od, ol), t, d in zip(outs, inps, grids, thr, dens):
        ref, st = oracle.apply_fixed(mra, P.tables(), f.nodes(), od, ol, t)
        c = o.nodes()[2]
        assert np.max(np.abs(ref - c)) / np.max(np.abs(ref)) < TOL_FP64
        total += st["term_apps"]
        E, E_ex = f.dot(o), cf.pair_energy(*d, mu=0.0)
        assert abs(E - E_ex) < 30 * prec * max(1.0, abs(E_ex))
    assert total == r64["term_apps"]
    print("batched plan: %d term-apps, mixed err %.2e" % (total, err_mx))


def test_helmholtz_batch_distinct_mu_energies(ctx):
    """One Helmholtz operator per orbital (mu_i), all in one batch; each
    <phi_i | G_mu_i phi_i> against the erfcx closed form."""
    mra = _mra(6)
    prec = 1e-5
    orbs = _orbitals(1005, 4, 3, 0.5, 15.0, 3.0)
    rng = workloads.SplitMix64(99)
    mus = [rng.uniform(0.5, 2.5) for _ in orbs]
    ops = [mw.HelmholtzOperator(mra, m, prec) for m in mus]
    inps = [mw.FunctionTree(mra).project_gaussians(prec, *o) for o in orbs]
    outs = [mw.FunctionTree(mra) for _ in orbs]
    mw.apply_batch([dict(prec=prec, out=o, oper=H, inp=f, abs_prec=True) for o, H, f in zip(outs, ops, inps)],
                   ctx=ctx)
    for o, f, orb, mu in zip(outs, inps, orbs, mus):
        E, E_ex = f.dot(o), cf.pair_energy(*orb, mu=mu)
        assert abs(E - E_ex) < 30 * prec * max(1.0, abs(E_ex)), (mu, E, E_ex)


def test_batch_argument_errors(ctx):
    mra = _mra(4)
    P = mw.PoissonOperator(mra, 1e-3)
    f = mw.FunctionTree(mra).project_gaussians(1e-3, [1.0], [5.0], np.zeros((1, 3)))
    o = mw.FunctionTree(mra)
    with pytest.raises(mw.MWGError):
        mw.apply_batch([dict(prec=1e-3, out=o, oper=P, inp=f), dict(prec=1e-3, out=o, oper=P, inp=f)], ctx=ctx)
    with pytest.raises(mw.MWGError):
        mw.apply_batch([dict(prec=1e-3, out=o, oper=P, inp=f), dict(prec=1e-3, out=f, oper=P, inp=o)], ctx=ctx)
    other = _mra(5)
    g = mw.FunctionTree(other).project_gaussians(1e-3, [1.0], [5.0], np.zeros((1, 3)))
    with pytest.raises(mw.MWGError):
        mw.apply_batch([dict(prec=1e-3, out=o, oper=P, inp=f),
                        dict(prec=1e-3, out=mw.FunctionTree(other), oper=P, inp=g)], ctx=ctx)
    assert mw.apply_batch([], ctx=ctx) == []
