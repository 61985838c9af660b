"""Benchmark: output nodes/s of the multiwavelet Helmholtz/Poisson apply.

Workload (default): the bound-state Helmholtz apply, mu = 1.2, on a 10-centre
Gaussian sum (exponents log-uniform in [0.5, 50], coefficients N(0,1), centres
uniform in [-3,3]^3), basis order k = 9, prec 1e-6, absolute precision as in
HelmholtzVector (src/scf_solver/HelmholtzVector.cpp:149).  The output grid
and per-node screening thresholds come from one adaptive apply and are then
frozen; a step is one apply over that frozen grid with the input tree, band
tables and operator tables already resident in HBM.

Prints one JSON line.  Roofline: algorithmic flops = term-applications that
passed screening x 6(k+1)^4 (counted by the kernel itself, equal to the
oracle's count), divided by the kernel's mean device time (HIP events on the
launch stream), against the MI355X FP64 peak.  cpu_baseline: the FP64 OpenMP
oracle (a restatement of the reference algorithm; MRCPP itself is not
available) timed on a bounded sample of the same output nodes on the host's
cores, in the same process.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "f16": 2500.0}
METRIC = "output nodes/s + % MFMA peak for Poisson/Helmholtz apply at prec 1e-6 vs plain FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(cfg_name, max_depth=16):
    import mrchem_amd as mw
    from mrchem_amd import workloads
    cfg = workloads.config(cfg_name)
    mra = mw.MultiResolutionAnalysis(order=cfg["order"], basis="interpolating", max_depth=max_depth,
                                     **workloads.WORLD)
    prec = cfg["prec"]
    t0 = time.time()
    if cfg["kind"] == "helmholtz":
        op = mw.HelmholtzOperator(mra, cfg["mu"][0], prec)
    else:
        op = mw.PoissonOperator(mra, prec)
    t_op = time.time() - t0
    c, a, R = cfg["densities"][0]
    t0 = time.time()
    rho = mw.FunctionTree(mra).project_gaussians(prec, c, a, R)
    t_proj = time.time() - t0
    return cfg, mra, op, rho, {"operator_build_s": t_op, "projection_s": t_proj}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import mrchem_amd as mw
    cfg, mra, op, rho, setup = build_workload(args.config)
    k = cfg["order"]
    log("workload %s: operator %d terms (%.2fs), input %d nodes (%.2fs)" %
        (cfg["name"], op.size(), setup["operator_build_s"], rho.size(), setup["projection_s"]))
    ctx = mw.Context.default()
    V = mw.FunctionTree(mra)
    t0 = time.time()
    mw.apply(cfg["prec"], V, op, rho, abs_prec=cfg["abs_prec"], ctx=ctx)
    t_adapt = time.time() - t0
    ast = ctx.last_stats()
    log("adaptive apply: %d output nodes in %.2fs (kernel %.1f ms, plan %.1f ms, %d launches)" %
        (V.size(), t_adapt, ast["kernel_ms"], ast["plan_ms"], ast["launches"]))
    thr = V.thresholds()
    od, ol, _ = V.nodes()
    # two plans on the same frozen grid and thresholds: plain FP64 and mixed
    V64 = V
    Vmx = mw.FunctionTree(mra).set_nodes(od, ol)
    t0 = time.time()
    plan64 = mw.ApplyPlan(V64, op, rho, thr, mode="fp64", ctx=ctx)
    planmx = mw.ApplyPlan(Vmx, op, rho, thr, mode="mixed", ctx=ctx)
    t_plan = time.time() - t0
    pinfo = plan64.info()
    log("frozen plans: %d out nodes, %d input slots (%d generated), %.2fs" %
        (pinfo["out_nodes"], pinfo["in_slots"], pinfo["generated"], t_plan))

    def timed(plan):
        for _ in range(args.warmup):
            plan.execute()
        runs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            runs.append(plan.execute())
        wall = time.perf_counter() - t0
        return runs, 1e3 * wall / args.steps

    runs64, ms64 = timed(plan64)
    runsmx, msmx = timed(planmx)
    n_out = pinfo["out_nodes"]
    apps = runs64[-1]["term_apps"]
    assert runsmx[-1]["term_apps"] == apps, "mixed and FP64 paths applied different term sets"
    lp_apps = runsmx[-1]["lp_term_apps"]
    fl = 6.0 * (k + 1) ** 4
    k64 = float(np.mean([r["ms"] for r in runs64]))
    kmx = float(np.mean([r["ms"] for r in runsmx]))
    kmx_lp = float(np.mean([r["ms_lp"] for r in runsmx]))
    kmx_64 = float(np.mean([r["ms_fp64"] for r in runsmx]))
    ach64 = apps * fl / (k64 * 1e-3) / 1e12
    ach_lp = lp_apps * fl / (kmx_lp * 1e-3) / 1e12
    log("fp64 : wall %.3f ms/step, kernel %.3f ms, %d term-apps, %.2f TFLOP/s" % (ms64, k64, apps, ach64))
    log("mixed: wall %.3f ms/step, kernel %.3f ms (fp64 part %.3f ms, fp32 part %.3f ms), %d of %d term-apps "
        "on FP32 (%.2f TFLOP/s)" % (msmx, kmx, kmx_64, kmx_lp, lp_apps, apps, ach_lp))

    # error of the mixed path against the FP64 path on the identical grid
    ns64 = plan64.download_ns()
    nsmx = planmx.download_ns()
    coef_err = float(np.max(np.abs(nsmx - ns64)) / np.max(np.abs(ns64)))
    plan64.finish()
    planmx.finish()
    c64 = V64.nodes()[2]
    cmx = Vmx.nodes()[2]
    tree_err = float(np.linalg.norm(cmx - c64) / np.linalg.norm(c64))
    log("mixed vs fp64: max coefficient error %.3e (relative to max), tree-norm error %.3e" % (coef_err, tree_err))

    # result check: energy against the closed form
    from oracle import closed_forms as cf
    c, a, R = cfg["densities"][0]
    E = rho.dot(V64)
    Emx = rho.dot(Vmx)
    mu = cfg["mu"][0] if cfg["kind"] == "helmholtz" else 0.0
    E_ex = cf.pair_energy(c, a, R, mu=mu)
    check = {"energy_fp64": E, "energy_mixed": Emx, "energy_closed_form": E_ex, "abs_err_fp64": abs(E - E_ex),
             "abs_err_mixed": abs(Emx - E_ex), "mixed_vs_fp64_max_coef_rel": coef_err,
             "mixed_vs_fp64_tree_rel": tree_err}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mra, op, rho, V64, thr, ns64, args.cpu_seconds, check)

    line = {
        "metric": METRIC,
        "value": n_out / (msmx * 1e-3),
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": msmx,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (mixed: FP64 for bound > %g thr)" % ctx_tau(ctx),
        "data": "synthetic (SplitMix64-seeded Gaussian sums)",
        "config": {"workload": cfg["name"], "order": k, "prec": cfg["prec"], "out_nodes": n_out,
                   "in_slots": pinfo["in_slots"], "terms": op.size(), "path": "mixed FP32/FP64 MFMA"},
        "roofline": {"bound": "mfma", "achieved": ach_lp, "peak": PEAK_TFLOPS["f32"], "unit": "TFLOP/s",
                     "frac": ach_lp / PEAK_TFLOPS["f32"], "traffic": None,
                     "kernel": "k_apply<%d, float>" % (k + 1)},
        "fp64_path": {"value": n_out / (ms64 * 1e-3), "ms_per_step": ms64, "kernel_ms": k64,
                      "roofline": {"achieved": ach64, "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                                   "frac": ach64 / PEAK_TFLOPS["f64"]}},
        "speedup_vs_fp64": ms64 / msmx,
        "term_apps": apps,
        "lp_term_apps": lp_apps,
        "check": check,
        "setup": dict(setup, adaptive_apply_s=t_adapt, plan_s=t_plan),
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


def ctx_tau(ctx):
    return 65536.0


def cpu_baseline(mra, op, rho, V, thr, ns_gpu, target_s, check):
    """FP64 OpenMP oracle on a bounded sample of the frozen output grid."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    oracle.set_threads(threads)
    od, ol, _ = V.nodes()
    n = len(od)
    tables = op.tables()
    inp = rho.nodes()
    rng = np.random.default_rng(12345)
    order = rng.permutation(n)
    m = min(n, 32)
    while True:
        idx = np.sort(order[:m])
        t0 = time.perf_counter()
        ref, st = oracle.apply_fixed(mra, tables, inp, od[idx], ol[idx], thr[idx], post=False)
        dt = time.perf_counter() - t0
        if m >= n or dt >= 0.5 * target_s:
            break
        m = min(n, max(m * 2, int(m * target_s / max(dt, 1e-3))))
    err = float(np.max(np.abs(ref - ns_gpu[idx])) / max(np.max(np.abs(ref)), 1e-300))
    check["sample_max_rel_err_vs_oracle"] = err
    log("cpu baseline: %d of %d nodes in %.2fs on %d threads; sample |gpu-oracle|/max = %.2e" %
        (m, n, dt, threads, err))
    return {"value": m / dt, "unit": "nodes/s", "cores": threads, "kind": "port",
            "sample": "%d of %d frozen-grid output nodes (random, seed 12345), FP64 OpenMP oracle, "
                      "no post transforms" % (m, n)}


if __name__ == "__main__":
    main()
