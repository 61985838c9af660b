"""Adaptive refinement (TreeBuilder + WaveletAdaptor) on the GPU against the
oracle's adaptive apply, and adaptive applies in the low-precision modes.

oracle.apply_adaptive grows its own output grid from the roots with the
reference rules: per refinement level the screening threshold
prec * sqrt(||out||^2 / M) from the running norm of the nodes computed so
far, and the split test ||d|| > prec * t * 2^{-(n+1)/2} (t = 1 with absolute
precision, else the running norm).  The GPU's adaptive mwg_apply must give
the same grid (identical node keys), the same per-node thresholds (1e-12
relative) and the same coefficients (1e-12 of the largest), on
- C1: the normalised Gaussian a = 10, Poisson, k = 7, prec 1e-5 relative to
  the density's norm (CoulombPotential's form);
- C2: the 10-centre Gaussian sum, Helmholtz mu = 1.2, k = 9, prec 1e-6
  absolute (HelmholtzVector's form).

The mixed and split16 modes run the same adaptive apply with their own
arithmetic; their grids may differ from FP64's only where a node's wavelet
norm sits at the split threshold, so they are compared as functions:
||V_mode - V_fp64|| (on the union grid, through mwg_add) within prec of
||V_fp64||, and their energies against the closed form.
"""
import numpy as np
import pytest

import mrchem_amd as mw
import oracle
from mrchem_amd import workloads
from oracle import closed_forms as cf

pytestmark = pytest.mark.gpu


def _load(name):
    cfg = workloads.config(name)
    mra = mw.MultiResolutionAnalysis(order=cfg["order"], basis="interpolating", max_depth=16, **workloads.WORLD)
    if cfg["kind"] == "helmholtz":
        op = mw.HelmholtzOperator(mra, cfg["mu"][0], cfg["prec"])
        mu = cfg["mu"][0]
    else:
        op = mw.PoissonOperator(mra, cfg["prec"])
        mu = 0.0
    d = cfg["densities"][0]
    rho = mw.FunctionTree(mra).project_gaussians(cfg["prec"], *d)
    prec = cfg["prec"] if cfg["abs_prec"] else cfg["prec"] / rho.norm()
    return cfg, mra, op, rho, prec, mu, d


def _keyed(depth, l):
    return [(int(a), int(b[0]), int(b[1]), int(b[2])) for a, b in zip(depth, l)]


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_adaptive_grid_equals_oracle(ctx, name):
    cfg, mra, op, rho, prec, mu, d = _load(name)
    V = mw.FunctionTree(mra)
    mw.apply(prec, V, op, rho, abs_prec=cfg["abs_prec"], ctx=ctx)
    gd, gl, gc = V.nodes()
    gthr = V.thresholds()
    od, ol, othr, oc, st = oracle.apply_adaptive(mra, op.tables(), rho.nodes(), prec, abs_prec=cfg["abs_prec"],
                                                 max_depth=mra.max_depth)
    gk, ok = _keyed(gd, gl), _keyed(od, ol)
    assert len(gk) == len(ok) and set(gk) == set(ok), (len(gk), len(ok))
    pos = {k: i for i, k in enumerate(ok)}
    perm = np.array([pos[k] for k in gk])
    thr_err = float(np.max(np.abs(gthr - othr[perm]) / np.maximum(othr[perm], 1e-300)))
    coef_err = float(np.max(np.abs(gc - oc[perm])) / np.max(np.abs(oc)))
    print("%s: %d nodes (depth <= %d) identical, threshold rel err %.2e, coefficient err %.2e, %d term-apps" %
          (name, len(gk), int(gd.max()), thr_err, coef_err, st["term_apps"]))
    assert thr_err < 1e-12
    assert coef_err < 1e-12
    E_ex = cf.pair_energy(*d, mu=mu)
    print("%s: energy %.12f, closed form %.12f, rel err %.2e (%.1f prec)" %
          (name, rho.dot(V), E_ex, abs(rho.dot(V) - E_ex) / abs(E_ex), abs(rho.dot(V) - E_ex) / abs(E_ex) /
           cfg["prec"]))


@pytest.mark.parametrize("mode", ["mixed", "split16"])
def test_adaptive_low_precision_on_c2(ctx, mode):
    cfg, mra, op, rho, prec, mu, d = _load("c2")
    V64 = mw.FunctionTree(mra)
    mw.apply(prec, V64, op, rho, abs_prec=True, ctx=ctx)
    Vm = mw.FunctionTree(mra)
    try:
        ctx.set_precision(mode)
        mw.apply(prec, Vm, op, rho, abs_prec=True, ctx=ctx)
        st = ctx.last_stats()
    finally:
        ctx.set_precision("fp64")
    same_grid = Vm.size() == V64.size() and np.array_equal(Vm.nodes()[1], V64.nodes()[1])
    diff = mw.add(-1.0, mw.FunctionTree(mra), [1.0, -1.0], [Vm, V64], ctx=ctx)
    rel = diff.norm() / V64.norm()
    E_ex = cf.pair_energy(*d, mu=mu)
    e64, em = rho.dot(V64), rho.dot(Vm)
    print("%s adaptive: %d nodes (fp64 %d, same grid: %s), ||V - V64|| / ||V64|| = %.2e, E err %.2e (fp64 %.2e), "
          "%d term-apps, kernels %.1f ms" % (mode, Vm.size(), V64.size(), same_grid, rel, abs(em - E_ex),
                                             abs(e64 - E_ex), st["term_apps"], st["kernel_ms"]))
    assert rel < cfg["prec"]
    assert abs(em - e64) < cfg["prec"] * max(1.0, abs(E_ex))
    assert abs(em - E_ex) < 30 * cfg["prec"] * max(1.0, abs(E_ex))
