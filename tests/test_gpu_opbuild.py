"""Operator tables built on the device (mwg_operators_create_device).

The device build must give the tables of the host build: the same kernel fit
(terms), the same stored half-width L for every (term, depth), and blocks and
norms equal to rounding (the sums run in another order).  An apply with the
device-built operator must then give the host-built operator's result, and
the Helmholtz energies of a batch of distinct mu must match their closed
forms.
"""
import time

import numpy as np
import pytest

import mrchem_amd as mw
from mrchem_amd import workloads
from oracle import closed_forms as cf

pytestmark = pytest.mark.gpu


def _mra(order, max_depth=16):
    return mw.MultiResolutionAnalysis(order=order, basis="interpolating", max_depth=max_depth, **workloads.WORLD)


def _host_op(mra, mu, prec):
    return mw.HelmholtzOperator(mra, mu, prec) if mu > 0 else mw.PoissonOperator(mra, prec)


def _compare_tables(th, td, kp1):
    """Max relative block and norm differences over (term, depth) entries
    with equal L, and the number of entries whose L differs."""
    M, nd = th["M"], th["ndepth"]
    assert (M, nd) == (td["M"], td["ndepth"])
    nn = kp1 * kp1
    nL = 0
    bmax = nmax = 0.0
    for i in range(M * nd):
        Lh, Ld = th["L"][i], td["L"][i]
        if Lh != Ld:
            nL += 1
            continue
        if Lh < 0:
            continue
        nb = 4 * (2 * Lh + 1)
        nh = th["norms"][th["noff"][i]:th["noff"][i] + nb]
        ndv = td["norms"][td["noff"][i]:td["noff"][i] + nb]
        bh = th["blocks"][th["boff"][i]:th["boff"][i] + nb * nn]
        bd = td["blocks"][td["boff"][i]:td["boff"][i] + nb * nn]
        bmax = max(bmax, np.max(np.abs(bh - bd)) / max(np.max(np.abs(bh)), 1e-300))
        nmax = max(nmax, np.max(np.abs(nh - ndv)) / max(np.max(nh), 1e-300))
    return bmax, nmax, nL


@pytest.mark.parametrize("order,mu,prec", [(4, 0.0, 1e-5), (7, 0.0, 1e-5), (9, 1.2, 1e-6), (11, 1.5, 1e-7)])
def test_device_tables_match_host(ctx, order, mu, prec):
    mra = _mra(order)
    host = _host_op(mra, mu, prec)
    t0 = time.time()
    (dev,) = mw.build_operators(mra, [mu], prec, ctx=ctx)
    t_dev = time.time() - t0
    ah, bh = host.terms()
    ad, bd = dev.terms()
    assert np.array_equal(ah, ad) and np.array_equal(bh, bd)
    th, td = host.tables(), dev.tables()
    bmax, nmax, nL = _compare_tables(th, td, order + 1)
    n_entries = int(np.sum(th["L"] >= 0))
    print("order %d mu %g: %d terms, %d stored (term, depth), L differs at %d, max block diff %.2e, "
          "norm diff %.2e, device build %.3f s" % (order, mu, len(ah), n_entries, nL, bmax, nmax, t_dev))
    # a stored half-width can flip only where a block norm sits at its floor
    assert nL <= max(1, n_entries // 500)
    assert bmax < 1e-12 and nmax < 1e-12


def test_device_operator_apply_matches_host(ctx):
    """Frozen-grid FP64 apply with the device-built Helmholtz operator against
    the host-built one, and the energy against its closed form."""
    mra = _mra(7)
    prec, mu = 1e-5, 1.1
    host = mw.HelmholtzOperator(mra, mu, prec)
    (dev,) = mw.build_operators(mra, [mu], prec, ctx=ctx)
    rng = workloads.SplitMix64(777)
    c, a, R = workloads.gaussian_sum(rng, 4, 0.5, 20.0, 2.0)
    rho = mw.FunctionTree(mra).project_gaussians(prec, c, a, R)
    V = mw.FunctionTree(mra)
    mw.apply(prec, V, host, rho, abs_prec=True, ctx=ctx)
    thr = V.thresholds()
    od, ol, _ = V.nodes()
    V2 = mw.FunctionTree(mra).set_nodes(od, ol)
    ph = mw.ApplyPlan(V, host, rho, thr, mode="fp64", ctx=ctx)
    pd = mw.ApplyPlan(V2, dev, rho, thr, mode="fp64", ctx=ctx)
    rh, rd = ph.execute(), pd.execute()
    nsh, nsd = ph.download_ns(), pd.download_ns()
    err = np.max(np.abs(nsh - nsd)) / np.max(np.abs(nsh))
    print("term-apps host %d device %d, max coef diff %.2e" % (rh["term_apps"], rd["term_apps"], err))
    assert abs(rh["term_apps"] - rd["term_apps"]) <= rh["term_apps"] // 10000
    assert err < 1e-11
    pd.finish()
    E = rho.dot(V2)
    Ecf = cf.pair_energy(c, a, R, mu=mu)
    assert abs(E - Ecf) < 30 * prec * max(1.0, abs(Ecf)), (E, Ecf)


def test_device_batch_of_mu(ctx):
    """Eight Helmholtz operators with distinct mu in one device build; two of
    them against their host builds."""
    mra = _mra(9)
    prec = 1e-6
    mus = np.linspace(0.5, 2.5, 8)
    t0 = time.time()
    ops = mw.build_operators(mra, mus, prec, ctx=ctx)
    t_dev = time.time() - t0
    t0 = time.time()
    hosts = [mw.HelmholtzOperator(mra, mus[i], prec) for i in (0, 7)]
    t_host = (time.time() - t0) / 2
    for h, i in zip(hosts, (0, 7)):
        bmax, nmax, nL = _compare_tables(h.tables(), ops[i].tables(), 10)
        assert nL <= 2 and bmax < 1e-12 and nmax < 1e-12
    print("8 operators on the device in %.3f s (host: %.3f s per operator)" % (t_dev, t_host))


def test_device_operator_other_context_is_an_error(ctx):
    mra = _mra(4, max_depth=8)
    (dev,) = mw.build_operators(mra, [0.0], 1e-4, ctx=ctx)
    other = mw.Context(0)
    f = mw.FunctionTree(mra).project_gaussians(1e-4, [1.0], [2.0], [[0.1, 0.0, 0.0]])
    with pytest.raises(mw.MWGError, match="another context"):
        mw.apply(1e-4, mw.FunctionTree(mra), dev, f, ctx=other)
    with pytest.raises(mw.MWGError):
        mw.build_operators(mra, [-1.0], 1e-4, ctx=ctx)
