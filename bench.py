"""Benchmark: output nodes/s of the multiwavelet Helmholtz/Poisson apply.

Default workload (c2): the bound-state Helmholtz apply, mu = 1.2, on a
10-centre Gaussian sum (exponents log-uniform in [0.5, 50], coefficients
N(0,1), centres uniform in [-3,3]^3), basis order k = 9, prec 1e-6, absolute
precision as in HelmholtzVector (src/scf_solver/HelmholtzVector.cpp:149).
--config c3 / c5 run the batched workloads: 136 exchange pair densities
(Poisson, k = 7) and 64 orbitals each with its own Helmholtz mu_i (k = 11).

The output grids and per-node screening thresholds come from one adaptive
(batched) apply and are then frozen; a step is one whole apply of every
entry over those grids with the input pools, band tables and operator tables
already resident in HBM: the per-node calculator, then the TopDown(add) and
BottomUp transforms and the node norms of the output, all on the device.
The adaptive apply itself (host refinement loop, the call MRChem makes) is
reported as its own rate, "adaptive_apply".

Prints one JSON line.  Roofline: algorithmic flops = term-applications that
passed screening x 6(k+1)^4 (counted by the kernel itself, equal to the
oracle's count), divided by the device time of the launch (HIP events on the
launch stream).  cpu_baseline: the FP64 OpenMP oracle (a restatement of the
reference algorithm; MRCPP itself is not available) timed on a bounded sample
of the same output nodes on the host's cores, in the same process.
"""
import argparse
import json
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


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summaries (profiles/*traffic*.md, lines
    'traffic_bytes <kernel> <bytes>'), or None when not measured."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        for line in open(path):
            if line.startswith("traffic_bytes ") and line[len("traffic_bytes "):].rsplit(" ", 1)[0] == kernel:
                return float(line.rsplit(" ", 1)[1])
    return None


def build_workload(cfg_name, max_depth=16):
    """Operators and projected input trees of one configuration (one entry
    per apply)."""
    import mrchem_amd as mw
    from mrchem_amd import workloads
    cfg = workloads.config(cfg_name)
    mra = mw.MultiResolutionAnalysis(order=cfg["order"], basis="interpolating", max_depth=max_depth,
                                     **workloads.WORLD)
    prec = cfg["prec"]
    dens = cfg["densities"]
    t0 = time.time()
    if cfg["kind"] == "helmholtz":
        mus = cfg["mu"] if len(cfg["mu"]) == len(dens) else [cfg["mu"][0]] * len(dens)
        cache = {}
        ops = []
        for m in mus:
            if m not in cache:
                cache[m] = mw.HelmholtzOperator(mra, m, prec)
            ops.append(cache[m])
    else:
        mus = [0.0] * len(dens)
        P = mw.PoissonOperator(mra, prec)
        ops = [P] * len(dens)
    t_op = time.time() - t0
    t0 = time.time()
    inps = [mw.FunctionTree(mra).project_gaussians(prec, c, a, R) for (c, a, R) in dens]
    t_proj = time.time() - t0
    return cfg, mra, ops, inps, mus, {"operator_build_s": t_op, "projection_s": t_proj}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="mixed,split16", help="low-precision modes timed beside plain FP64")
    ap.add_argument("--mode", default="split16", help="the mode the headline value is quoted for")
    ap.add_argument("--tau", type=float, default=None, help="FP64 class cut (bound > tau x threshold)")
    ap.add_argument("--no-graph", action="store_true", help="time eager launches instead of graph replays")
    ap.add_argument("--no-tune", action="store_true", help="keep the default knobs of the low-precision plans")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import mrchem_amd as mw
    cfg, mra, ops, inps, mus, setup = build_workload(args.config)
    k = cfg["order"]
    n_app = len(inps)
    log("workload %s: %d applies, operator(s) %d terms (%.2fs), inputs %d nodes (%.2fs)" %
        (cfg["name"], n_app, ops[0].size(), setup["operator_build_s"], sum(f.size() for f in inps),
         setup["projection_s"]))
    ctx = mw.Context.default()
    if args.tau is not None:
        ctx.set_precision("fp64", args.tau)
    # the same operators built on the device (one build for all distinct mu);
    # the applies below use the host-built ones
    umus = sorted(set(mus))
    mw.build_operators(mra, umus[:1], cfg["prec"], ctx=ctx)
    t0 = time.time()
    dev_ops = mw.build_operators(mra, umus, cfg["prec"], ctx=ctx)
    setup["operator_build_device_s"] = time.time() - t0
    del dev_ops
    log("operators built on the device: %d in %.3fs (host: %.2fs)" % (len(umus), setup["operator_build_device_s"],
                                                                      setup["operator_build_s"]))
    if cfg.get("orbitals"):
        # pair density rho_01 = phi_0 phi_1 formed on the device (mwg_multiply)
        # against the projection of the exact product (the apply's input 1)
        phi = [mw.FunctionTree(mra).project_gaussians(cfg["prec"], *o) for o in cfg["orbitals"][:2]]
        reps = []
        for _ in range(3):
            t0 = time.time()
            rho = mw.multiply(cfg["prec"], mw.FunctionTree(mra), phi[0], phi[1], abs_prec=True, ctx=ctx)
            reps.append(time.time() - t0)
        ref = inps[1]
        diff = float(np.sqrt(max(rho.norm() ** 2 + ref.norm() ** 2 - 2.0 * rho.dot(ref), 0.0)))
        setup["pair_density_device"] = {"s": min(reps), "out_nodes": rho.size(),
                                        "in_nodes": [phi[0].size(), phi[1].size()], "diff_norm": diff,
                                        "norm": ref.norm()}
        if not args.no_cpu:
            # the same node products by the oracle's restatement on the host
            # cores, at every node of the device's output grid
            import oracle
            od, ol, _ = rho.nodes()
            t0 = time.time()
            oracle.multiply(mra, phi[0].nodes(), phi[1].nodes(), od, ol)
            t_cpu = time.time() - t0
            setup["pair_density_device"]["cpu_oracle_s"] = t_cpu
            setup["pair_density_device"]["cpu_threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            log("pair density on the host (oracle, all %d nodes): %.3fs" % (len(od), t_cpu))
        log("pair density phi_0 phi_1 on the device: %d nodes (inputs %d, %d) in %.3fs, |rho - exact| %.2e "
            "(|exact| %.3e)" % (rho.size(), phi[0].size(), phi[1].size(), min(reps), diff, ref.norm()))
        # linear combination of 16 pair densities on their union grid (the
        # exchange sums, mwg_add); checked by the norm of the sum against the
        # pairwise inner products
        m = 16
        cs = np.linspace(1.0, -1.0, m)
        reps = []
        for _ in range(3):
            t0 = time.time()
            lc = mw.add(-1.0, mw.FunctionTree(mra), cs, inps[:m], ctx=ctx)
            reps.append(time.time() - t0)
        gram = np.array([[inps[i].dot(inps[j]) for j in range(m)] for i in range(m)])
        sq = float(cs @ gram @ cs)
        setup["linear_combination_device"] = {"s": min(reps), "inputs": m, "out_nodes": lc.size(),
                                              "in_nodes": int(sum(f.size() for f in inps[:m])),
                                              "norm_rel_err": abs(lc.norm() ** 2 - sq) / sq}
        log("linear combination of %d pair densities on the device: %d nodes (inputs %d nodes) in %.3fs, "
            "norm identity %.1e" % (m, lc.size(), setup["linear_combination_device"]["in_nodes"], min(reps),
                                     setup["linear_combination_device"]["norm_rel_err"]))
    prec = cfg["prec"]

    def adaptive(resident):
        """One adaptive mwg_apply_batch from host trees, or with inputs and
        outputs resident on the device (the result stays there)."""
        outs = [mw.FunctionTree(mra) for _ in inps]
        for t in (outs + inps) if resident else ():
            t.set_resident(ctx)
        entries = [dict(prec=prec if cfg["abs_prec"] else prec / f.norm(), out=o, oper=op, inp=f,
                        abs_prec=cfg["abs_prec"]) for o, op, f in zip(outs, ops, inps)]
        t0 = time.time()
        mw.apply_batch(entries, ctx=ctx)
        dt = time.time() - t0
        return outs, dt, ctx.last_stats()

    adaptive(False)  # first call: code objects, allocator and operator tables warm up
    outs, t_adapt, ast = adaptive(False)
    n_out = sum(o.size() for o in outs)
    log("adaptive apply: %d output nodes in %.1f ms (kernel %.1f ms, plan %.1f ms = bands %.1f + pools %.1f + "
        "counts %.1f, %d launches; device plan create %.1f + run %.1f of %.1f ms, split checks %.1f ms, post "
        "transforms + download %.1f ms)" %
        (n_out, 1e3 * t_adapt, ast["kernel_ms"], ast["plan_ms"], ast["plan_grid_ms"], ast["plan_pool_ms"],
         ast["plan_counts_ms"], ast["launches"], ast["device_create_ms"], ast["device_run_ms"], ast["device_ms"],
         ast["split_ms"], ast["post_ms"]))
    adaptive(True)  # uploads the inputs once
    routs, t_res, rst = adaptive(True)
    log("adaptive apply, trees resident on the device: %.1f ms (plan %.1f ms, device %.1f ms, post %.1f ms)" %
        (1e3 * t_res, rst["plan_ms"], rst["device_ms"], rst["post_ms"]))
    assert sum(o.size() for o in routs) == n_out
    for t in routs + inps:
        t.set_resident(ctx, False)
    thr = [o.thresholds() for o in outs]
    grids = [o.nodes()[:2] for o in outs]
    # plans on the same frozen grids and thresholds: plain FP64 and the
    # low-precision modes (FP64 for bound > tau thr, the rest in FP32 or as
    # FP16 split products)
    modes = ["fp64"] + [m for m in args.modes.split(",") if m != "fp64"]
    t0 = time.time()
    plans = {}
    trees = {}
    prec = cfg["prec"]
    for m in modes:
        trees[m] = outs if m == "fp64" else [mw.FunctionTree(mra).set_nodes(*g) for g in grids]
        plans[m] = mw.ApplyPlan(trees[m], ops, inps, thr, mode=m, ctx=ctx)
    t_plan = time.time() - t0
    # the autotuner: FP64-class cut, single-rounding cut and queue order of
    # each low-precision plan, within 0.1 prec of the FP64 result
    tuning = {}
    if not args.no_tune:
        for m in modes[1:]:
            t0 = time.time()
            tuning[m] = plans[m].tune(0.1 * prec, reps=3)
            tuning[m]["tune_s"] = time.time() - t0
            log("tuned %s: tau %g, single-rounding ratio %g, queue order %d: %.3f -> %.3f ms (%.2fx), error %.2e "
                "(%d settings, %.1fs)" % (m, tuning[m]["tau"], tuning[m]["single_ratio"], tuning[m]["order"],
                                         tuning[m]["start_ms"], tuning[m]["ms"], tuning[m]["speedup"],
                                         tuning[m]["rel_err"], tuning[m]["tried"], tuning[m]["tune_s"]))
    pinfo = plans["fp64"].info()
    log("frozen plans: %d out nodes, %d input slots (%d generated), %d segments, band half-width <= %d, %.2fs" %
        (pinfo["out_nodes"], pinfo["in_slots"], pinfo["generated"], pinfo["segments"], pinfo["band_w"], t_plan))

    def timed(plan):
        """Warmup steps, eager steps (HIP events around each launch: the
        per-launch device times), then the timed steps, each one replay of
        the captured graph of the whole apply (or eager with --no-graph)."""
        for _ in range(args.warmup):
            plan.apply()
        eager = [plan.apply() for _ in range(max(3, args.steps // 4))]
        if not args.no_graph:
            plan.use_graph(True)
            plan.apply()  # capture and first replay
        runs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            runs.append(plan.apply())
        wall = time.perf_counter() - t0
        plan.use_graph(False)
        return runs, eager, 1e3 * wall / args.steps

    fl = 6.0 * (k + 1) ** 4
    res = {}
    for m in modes:
        runs, eager, wall = timed(plans[m])
        r = {"wall_ms": wall, "kernel_ms": float(np.mean([x["ms"] for x in runs])), "term_apps": runs[-1]["term_apps"],
             "lp_term_apps": runs[-1]["lp_term_apps"], "ms_fp64": float(np.mean([x["ms_fp64"] for x in eager])),
             "ms_lp": float(np.mean([x["ms_lp"] for x in eager])),
             "post_ms": float(np.mean([x["post_ms"] for x in eager])),
             "eager_ms": float(np.mean([x["ms"] for x in eager])), "graph": not args.no_graph}
        res[m] = r
    apps = res["fp64"]["term_apps"]
    assert apps == pinfo["host_term_apps"], "kernel and host screening counts differ"
    for m in modes:
        assert res[m]["term_apps"] == apps, "%s applied a different term set" % m
    # the FP64 calculator alone (its one launch), for its roofline
    calc64 = res["fp64"]["ms_fp64"]
    ach64 = apps * fl / (calc64 * 1e-3) / 1e12
    log("fp64   : wall %.3f ms/step, device %.3f ms (calculator %.3f ms, post transforms + norms %.3f ms), %d "
        "term-apps, %.2f TFLOP/s" % (res["fp64"]["wall_ms"], res["fp64"]["kernel_ms"], calc64,
                                      res["fp64"]["post_ms"], apps, ach64))
    # raw non-standard blocks for the checks: one calculator run per mode
    for m in modes:
        plans[m].execute()
    ns64 = plans["fp64"].download_ns()
    for m in modes[1:]:
        r = res[m]
        r["lp_tflops"] = r["lp_term_apps"] * fl / (r["ms_lp"] * 1e-3) / 1e12 if r["ms_lp"] > 0 else 0.0
        ns = plans[m].download_ns()
        r["coef_err"] = float(np.max(np.abs(ns - ns64)) / np.max(np.abs(ns64)))
        log("%-7s: wall %.3f ms/step, device %.3f ms (fp64 part %.3f ms, low-precision part %.3f ms, post "
            "transforms + norms %.3f ms), %d of %d term-apps low-precision (%.2f TFLOP/s algorithmic), max coef err "
            "vs fp64 %.3e" % (m, r["wall_ms"], r["kernel_ms"], r["ms_fp64"], r["ms_lp"], r["post_ms"],
                              r["lp_term_apps"], apps, r["lp_tflops"], r["coef_err"]))
    for m in modes:
        plans[m].finish()
    c64 = np.concatenate([o.nodes()[2] for o in outs])
    for m in modes[1:]:
        cm = np.concatenate([o.nodes()[2] for o in trees[m]])
        res[m]["tree_err"] = float(np.linalg.norm(cm - c64) / np.linalg.norm(c64))
    head = args.mode if args.mode in res else modes[-1]
    hr = res[head]

    # result check: energies against the closed forms
    from oracle import closed_forms as cf
    e_err = {m: 0.0 for m in modes}
    e_max = 0.0
    for a_i, (f, d, mu) in enumerate(zip(inps, cfg["densities"], mus)):
        E_ex = cf.pair_energy(*d, mu=mu)
        e_max = max(e_max, abs(E_ex))
        for m in modes:
            e_err[m] = max(e_err[m], abs(f.dot(trees[m][a_i]) - E_ex))
    check = {"energy_max_abs_err": e_err, "energy_max_abs": e_max,
             "vs_fp64_max_coef_rel": {m: res[m]["coef_err"] for m in modes[1:]},
             "vs_fp64_tree_rel": {m: res[m]["tree_err"] for m in modes[1:]}}
    if n_app == 1:
        check["energy_closed_form"] = cf.pair_energy(*cfg["densities"][0], mu=mus[0])
    log("energies vs closed form: max abs err %s (max |E| %.3e)" %
        (", ".join("%s %.3e" % (m, e_err[m]) for m in modes), e_max))
    msmx, ms64 = hr["wall_ms"], res["fp64"]["wall_ms"]
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mra, ops, inps, outs, thr, ns64, args.cpu_seconds, check)

    peak_key = "f16" if head == "split16" else "f32"
    kname = ("k_apply_h<%d>" if head == "split16" else "k_apply<%d, float>") % (k + 1)
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
        "dtype": ("f16x3+f64 (FP16 hi/lo split products, FP32 accumulate; FP64 for bound > %g thr)" if head == "split16"
                  else "f32+f64 (mixed: FP64 for bound > %g thr)") % plans[head].info()["tau"],
        "data": "synthetic (SplitMix64-seeded Gaussian sums)",
        "config": {"workload": cfg["name"], "order": k, "prec": prec, "applies": n_app, "out_nodes": n_out,
                   "in_slots": pinfo["in_slots"], "terms": ops[0].size(), "path": head},
        "roofline": {"bound": "mfma", "achieved": hr["lp_tflops"], "peak": PEAK_TFLOPS[peak_key], "unit": "TFLOP/s",
                     "frac": hr["lp_tflops"] / PEAK_TFLOPS[peak_key], "traffic": measured_traffic(kname),
                     "kernel": kname},
        "fp64_path": {"value": n_out / (ms64 * 1e-3), "ms_per_step": ms64, "kernel_ms": res["fp64"]["kernel_ms"],
                      "roofline": {"achieved": ach64, "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
                                   "frac": ach64 / PEAK_TFLOPS["f64"], "kernel": "k_apply<%d, double>" % (k + 1)}},
        "adaptive_apply": {"value": n_out / t_adapt, "unit": "nodes/s", "wall_ms": 1e3 * t_adapt,
                           "note": "the adaptive drop-in call (mwg_apply_batch, FP64) that fixed the grid, second "
                                   "call of the process: host refinement loop, per-level plans and launches, post "
                                   "transforms, download into host trees",
                           "stages_ms": {k: ast[k] for k in ("kernel_ms", "plan_ms", "plan_grid_ms", "plan_pool_ms",
                                                             "plan_counts_ms", "device_ms", "device_create_ms",
                                                             "device_run_ms", "split_ms", "post_ms", "wall_ms")},
                           "resident": {"value": n_out / t_res, "wall_ms": 1e3 * t_res,
                                        "note": "input and output trees resident on the device "
                                                "(mwg_tree_set_resident): no block upload or download",
                                        "stages_ms": {k: rst[k] for k in ("kernel_ms", "plan_ms", "device_ms",
                                                                          "split_ms", "post_ms", "wall_ms")}}},
        "modes": {m: dict(res[m], value=n_out / (res[m]["wall_ms"] * 1e-3)) for m in modes},
        "speedup_vs_fp64": ms64 / msmx,
        "term_apps": apps,
        "lp_term_apps": hr["lp_term_apps"],
        "check": check,
        "tuning": tuning,
        "setup": dict(setup, adaptive_apply_s=t_adapt, plan_s=t_plan,
                      adaptive_apply_ms={k: ast[k] for k in ("kernel_ms", "plan_ms", "device_ms", "split_ms",
                                                             "post_ms", "wall_ms")}),
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


def host_cpu():
    """(model name, physical cores per socket, sockets) from lscpu, and the
    CPUs this process may run on."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    model = info.get("Model name") or cpu_model()
    try:
        per_socket = int(info.get("Core(s) per socket", "0"))
        sockets = int(info.get("Socket(s)", "1"))
    except ValueError:
        per_socket, sockets = 0, 1
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, per_socket, sockets, usable


def cpu_baseline(mra, ops, inps, outs, thr, ns_gpu, target_s, check):
    """FP64 OpenMP oracle on a bounded sample of the frozen output grids:
    whole applies for a batch (in a seeded random order), random nodes of the
    one grid for a single apply.  Threads: the physical cores of one socket,
    or fewer when this process may run on fewer CPUs (the GPU box gives each
    GPU a share of its host CPUs)."""
    import oracle
    model, per_socket, sockets, usable = host_cpu()
    threads = min(per_socket, usable) if per_socket > 0 else usable
    oracle.set_threads(threads)
    rng = np.random.default_rng(12345)
    offs = np.cumsum([0] + [o.size() for o in outs])
    n_tot = int(offs[-1])
    if len(outs) == 1:
        od, ol, _ = outs[0].nodes()
        n = len(od)
        order = rng.permutation(n)
        m = min(n, 32)
        tables = ops[0].tables()
        inp = inps[0].nodes()
        while True:
            idx = np.sort(order[:m])
            t0 = time.perf_counter()
            ref, _ = oracle.apply_fixed(mra, tables, inp, od[idx], ol[idx], thr[0][idx], post=False)
            dt = time.perf_counter() - t0
            if m >= n or dt >= 0.5 * target_s:
                break
            m = min(n, max(m * 2, int(m * target_s / max(dt, 1e-3))))
        err = float(np.max(np.abs(ref - ns_gpu[idx])) / max(np.max(np.abs(ref)), 1e-300))
        sample = "%d of %d frozen-grid output nodes (random, seed 12345)" % (m, n)
        done = m
        # one thread on a smaller sample of the same nodes
        m1 = max(8, m // 32)
        idx1 = np.sort(order[:m1])
        oracle.set_threads(1)
        t0 = time.perf_counter()
        oracle.apply_fixed(mra, tables, inp, od[idx1], ol[idx1], thr[0][idx1], post=False)
        one_thread = m1 / (time.perf_counter() - t0)
        oracle.set_threads(threads)
    else:
        order = rng.permutation(len(outs))
        done, dt, err, used = 0, 0.0, 0.0, 0
        tabs = {}
        for a in order:
            od, ol, _ = outs[a].nodes()
            key = id(ops[a])
            if key not in tabs:
                tabs[key] = ops[a].tables()
            t0 = time.perf_counter()
            ref, _ = oracle.apply_fixed(mra, tabs[key], inps[a].nodes(), od, ol, thr[a], post=False)
            dt += time.perf_counter() - t0
            g = ns_gpu[offs[a]:offs[a + 1]]
            err = max(err, float(np.max(np.abs(ref - g)) / max(np.max(np.abs(ref)), 1e-300)))
            done += len(od)
            used += 1
            if dt >= target_s:
                break
        m = done
        sample = "%d of %d applies (%d of %d output nodes, random order, seed 12345)" % (used, len(outs), done, n_tot)
    check["sample_max_rel_err_vs_oracle"] = err
    log("cpu baseline: %d of %d nodes in %.2fs on %d threads; sample |gpu-oracle|/max = %.2e" %
        (m, n_tot, dt, threads, err))
    res = {"value": m / dt, "unit": "nodes/s", "cores": threads, "kind": "port",
           "sample": sample + ", FP64 OpenMP oracle, no post transforms", "cpu_model": model,
           "cores_per_socket": per_socket, "sockets": sockets, "cpus_usable": usable}
    if len(outs) == 1:
        res["one_thread_value"] = one_thread
        log("cpu baseline, 1 thread: %.1f nodes/s" % one_thread)
    return res


def cpu_model():
    """Model name of the host CPU (from /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
