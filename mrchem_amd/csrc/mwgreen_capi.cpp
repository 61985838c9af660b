// C ABI of the multiwavelet Green's-function apply and its host drivers.
//
// mwg_apply restates mrcpp::apply (treebuilders/apply.cpp upstream, fetched by
// external/upstream/fetch_mrcpp.cmake:36-43; callers
// src/qmoperators/two_electron/CoulombPotential.cpp:131,
// src/qmoperators/two_electron/ExchangePotential.cpp:200,
// src/scf_solver/HelmholtzVector.cpp:149): the output grid grows level by
// level on the host (TreeBuilder + WaveletAdaptor), each level's nodes are
// computed by one HIP launch, and the result is finished by the TopDown(add)
// and BottomUp transforms.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <set>
#include <string>
#include <vector>

#include "../../include/mwgreen.h"
#include "mw_device.hpp"
#include "mw_internal.hpp"

using namespace mwg;

// Knobs of a frozen-grid plan chosen by mwg_plan_tune.
struct TunedKnobs {
    double tau = 65536.0, single_ratio = 16.0;
    int order = 0;
};
struct mwg_context_s {
    DeviceCtx *dev = nullptr;
    // tuned knobs by (k + 1, terms of the first operator, widest band, mode):
    // new plans with the same key start from them
    std::map<std::tuple<int, int, int, int>, TunedKnobs> tuned;
    std::vector<double> last_stats = std::vector<double>(8, 0.0);
    int mode = MWG_MODE_FP64;
    double tau = 65536.0;    // mixed mode: FP64 for terms with bound > tau * thr
    std::atomic<int> plan_refs{0};  // live plans on this context
    std::string last_err;           // message of the last failed call on this context
    std::atomic<int> resident_refs{0};  // trees whose blocks are resident on this context's GPU
    // bytes of tree blocks moved between host and device by calls on this
    // context (mwg_context_transfer_bytes)
    std::atomic<long long> h2d_bytes{0}, d2h_bytes{0};
};
struct mwg_operator_s {
    Operator op;
    OperFlat flat;
    std::shared_ptr<MRA> mra_keepalive;
    mwg_context dev_ctx = nullptr;  // device-built: the context holding the tables
    std::atomic<int> plan_refs{0};  // live plans holding pointers to its device tables
};
struct mwg_tree_s {
    Tree tree;
    std::vector<double> apply_thr;
    std::atomic<int> plan_refs{0};  // live plans that write their result into this tree
    // Residency (mwg_tree_set_resident): the blocks in HBM of dctx's GPU,
    // tree.size() x 8 (k+1)^3 doubles in tree order at dblocks (dcap
    // allocated).  host_valid: tree.coefs holds the blocks; dev_valid:
    // dblocks does.  At least one of them holds; a resident tree written by a
    // device call on dctx holds only the device copy until host code reads it.
    mwg_context dctx = nullptr;
    double *dblocks = nullptr;
    size_t dcap = 0;
    bool host_valid = true, dev_valid = false;
    // per-node squared norms of the scaling part (s2) and of the wavelet
    // parts (w2) of the current blocks, when norms_valid
    std::vector<double> s2, w2;
    bool norms_valid = false;
    // the 8 component norms of each node, when cn8_valid (apply inputs)
    std::vector<double> cn8;
    bool cn8_valid = false;
};
struct mwg_plan_s {
    mwg_context ctx = nullptr;
    DevicePlan *dev = nullptr;
    std::vector<mwg_tree> outs;           // per apply of the batch
    std::vector<mwg_operator> ops;        // per apply of the batch
    std::vector<std::vector<NodeKey>> keys;  // output grids at plan creation
    std::vector<std::vector<double>> thr;
    std::vector<int> n_out;
    int total_out = 0;
    int mode = MWG_MODE_FP64;
    double tau = 65536.0;  // FP64-class cut of this plan (the context's at creation, or tuned)
    std::tuple<int, int, int, int> tune_key;
    double plan_ms = 0.0, slots = 0.0, gen = 0.0, host_apps = 0.0, segs = 0.0, band_w = 0.0;
    // the last execution ran the post transforms in place (mwg_plan_apply):
    // the device holds the finished blocks and their norms
    bool post_done = false;
    // executions replay a captured hipGraph of the step (mwg_plan_use_graph)
    bool use_graph = false;
    // The plan holds device pointers into its operators' resident tables and
    // writes into its output trees: each of them, and the context, counts the
    // plan until it is destroyed, and refuses to be destroyed meanwhile.
    void hold() {
        ctx->plan_refs++;
        for (auto *o : outs) o->plan_refs++;
        for (auto *o : ops) o->plan_refs++;
    }
    ~mwg_plan_s() {
        if (!held) return;
        ctx->plan_refs--;
        for (auto *o : outs) o->plan_refs--;
        for (auto *o : ops) o->plan_refs--;
    }
    bool held = false;
};
struct mwg_mra_s {
    MRA mra;
};

namespace {
thread_local std::string g_err;
std::mutex g_ctx_mtx;
std::set<mwg_context_s *> g_contexts;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

void ensure_norms(mwg_tree t);
// A failed device step inside a call (rc -1: an argument the device code
// refused, else a HIP error).
struct DeviceFailure : std::runtime_error {
    int rc;
    DeviceFailure(int r, const std::string &m) : std::runtime_error(m), rc(r) {}
};
void dev_check(int rc, const std::string &err) {
    if (rc) throw DeviceFailure(rc, err);
}

template <typename F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const std::invalid_argument &e) {
        return fail(MWG_ERR_ARG, e.what());
    } catch (const DeviceFailure &e) {
        return fail(e.rc == -1 ? MWG_ERR_ARG : MWG_ERR_DEVICE, e.what());
    } catch (const std::exception &e) {
        return fail(MWG_ERR_RUNTIME, e.what());
    }
}

// Calls on a context (or a plan of it) also keep their failure message in the
// context: mwg_context_last_error.
int on_ctx(mwg_context ctx, int rc) {
    if (rc != MWG_OK && ctx) ctx->last_err = g_err;
    return rc;
}

template <typename F>
int guarded_ctx(mwg_context ctx, F &&f) {
    return on_ctx(ctx, guarded(std::forward<F>(f)));
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A prec tree: its grid and the squared norms of each node's scaling and
// wavelet parts.
struct PrecTree {
    const Tree *t;
    const std::vector<double> *s2, *w2;
};

// Norm of the node at key in a prec tree, or of its deepest ancestor scaled
// by the volume ratio when the tree is coarser there.
double prec_tree_norm(const PrecTree &p, const NodeKey &k) {
    const Tree &t = *p.t;
    NodeKey a = k;
    int up = 0;
    while (true) {
        int i = t.find(a.depth, a.l);
        if (i >= 0) return std::sqrt((*p.s2)[i] + (*p.w2)[i]) * std::pow(2.0, -1.5 * up);
        if (a.depth == 0) return 0.0;
        a = NodeKey{a.depth - 1, {a.l[0] >> 1, a.l[1] >> 1, a.l[2] >> 1}};
        up++;
    }
}

double prec_factor(const std::vector<PrecTree> &pt, const NodeKey &k) {
    if (pt.empty()) return 1.0;
    double pn = 1.0;
    for (const PrecTree &t : pt) pn = std::max(pn, prec_tree_norm(t, k));
    return 1.0 / pn;
}

void check_compatible(const Tree &out, const Operator &op, const Tree &inp) {
    if (out.mra != op.mra || inp.mra != op.mra)
        throw std::invalid_argument("apply: trees and operator were built on different MRAs");
    // the device kernels are instantiated for k + 1 = 2..12
    if (op.mra->kp1() > 12)
        throw std::invalid_argument("apply: the device kernels take k + 1 <= 12 (order <= 11), this MRA has k + 1 = " +
                                    std::to_string(op.mra->kp1()));
    if (inp.size() == 0) throw std::invalid_argument("apply: input tree is empty");
}

int dev_fail(int rc, const std::string &err) { return fail(rc == -1 ? MWG_ERR_ARG : MWG_ERR_DEVICE, err); }

// ---- Device-resident trees ----

size_t tree_values(const mwg_tree_s *t) { return (size_t)t->tree.size() * t->tree.mra->block(); }

// The host copy of t's blocks, downloaded when only the device holds them.
void ensure_host(mwg_tree t) {
    if (t->host_valid) return;
    const size_t n = tree_values(t);
    t->tree.structure_only = false;
    t->tree.coefs.resize(n);
    std::string err;
    dev_check(dev_download(t->dctx->dev, t->dblocks, n, t->tree.coefs.data(), err), err);
    t->dctx->d2h_bytes += (long long)(sizeof(double) * n);
    t->host_valid = true;
}

// Host code changed t's blocks or grid: the device copy and the norms are
// stale.
void host_changed(mwg_tree t) {
    t->host_valid = true;
    t->dev_valid = false;
    t->norms_valid = false;
    t->cn8_valid = false;
}

// The 8 component norms of every node of t (on the device for a tree whose
// blocks only the device holds).
const double *ensure_cnorms(mwg_tree t) {
    if (t->cn8_valid) return t->cn8.data();
    const Tree &T = t->tree;
    const int n = T.size();
    t->cn8.assign((size_t)8 * n, 0.0);
    if (!t->host_valid) {
        std::string err;
        dev_check(dev_component_norms(t->dctx->dev, T.mra->kp1(), t->dblocks, n, t->cn8.data(), err), err);
    } else {
#pragma omp parallel for num_threads(team_size()) schedule(static)
        for (int i = 0; i < n; i++) component_norms(*T.mra, T.node(i), t->cn8.data() + (size_t)8 * i);
    }
    t->cn8_valid = true;
    return t->cn8.data();
}

// t's blocks in device memory of ctx: the resident copy (uploaded first when
// stale), or else an upload into *tmp, which the caller frees (dev_free).
const double *device_blocks(mwg_context ctx, mwg_tree t, double **tmp) {
    *tmp = nullptr;
    const size_t n = tree_values(t);
    std::string err;
    if (t->dctx == ctx) {
        if (!t->dev_valid) {
            if (t->dcap < n) {
                dev_free(ctx->dev, t->dblocks);
                t->dblocks = nullptr;
                t->dcap = 0;
                dev_check(dev_alloc(ctx->dev, n, &t->dblocks, err), err);
                t->dcap = n;
            }
            dev_check(dev_upload_into(ctx->dev, t->tree.coefs.data(), n, t->dblocks, err), err);
            ctx->h2d_bytes += (long long)(sizeof(double) * n);
            t->dev_valid = true;
        }
        return t->dblocks;
    }
    ensure_host(t);
    dev_check(dev_upload(ctx->dev, t->tree.coefs.data(), n, tmp, err), err);
    ctx->h2d_bytes += (long long)(sizeof(double) * n);
    return *tmp;
}

// Frees the temporary device copies of a call's input trees.
struct TempBlocks {
    DeviceCtx *c;
    std::vector<double *> p;
    ~TempBlocks() {
        for (double *q : p) dev_free(c, q);
    }
    const double *of(mwg_context ctx, mwg_tree t) {
        double *tmp = nullptr;
        const double *b = device_blocks(ctx, t, &tmp);
        if (tmp) p.push_back(tmp);
        return b;
    }
};

// The per-node norms of t's blocks (computed on the device for a tree whose
// blocks only the device holds).
void ensure_norms(mwg_tree t) {
    if (t->norms_valid) return;
    const Tree &T = t->tree;
    const int n = T.size(), n3 = T.mra->kp1_3(), blk = T.mra->block();
    t->s2.assign(n, 0.0);
    t->w2.assign(n, 0.0);
    if (!t->host_valid) {
        std::string err;
        dev_check(dev_node_norms(t->dctx->dev, T.mra->kp1(), t->dblocks, n, t->s2.data(), t->w2.data(), err), err);
    } else {
#pragma omp parallel for num_threads(team_size()) schedule(static)
        for (int i = 0; i < n; i++) {
            const double *b = T.node(i);
            double a = 0.0, w = 0.0;
            for (int e = 0; e < n3; e++) a += b[e] * b[e];
            for (int e = n3; e < blk; e++) w += b[e] * b[e];
            t->s2[i] = a;
            t->w2[i] = w;
        }
    }
    t->norms_valid = true;
}

// A result computed on the device by a call on ctx: t's blocks (tree order,
// tree_values(t) doubles, owned by the call until handed over here) and their
// norms.  Kept as t's resident copy when t is resident on ctx, else
// downloaded into the host tree and freed.
void take_device_result(mwg_context ctx, mwg_tree t, double *blocks, std::vector<double> &&s2,
                        std::vector<double> &&w2) {
    const size_t n = tree_values(t);
    Tree &T = t->tree;
    T.structure_only = false;
    if (t->dctx == ctx) {
        if (t->dblocks && t->dblocks != blocks) dev_free(ctx->dev, t->dblocks);
        t->dblocks = blocks;
        t->dcap = n;
        t->dev_valid = true;
        t->host_valid = false;
        T.coefs.clear();
    } else {
        T.coefs.resize(n);
        std::string err;
        const int rc = dev_download(ctx->dev, blocks, n, T.coefs.data(), err);
        dev_free(ctx->dev, blocks);
        dev_check(rc, err);
        ctx->d2h_bytes += (long long)(sizeof(double) * n);
        t->host_valid = true;
        t->dev_valid = false;
    }
    double sq = 0.0;
    for (int i = 0; i < T.size(); i++) sq += w2[i] + (T.parent[i] < 0 ? s2[i] : 0.0);
    T.sqnorm = sq;
    t->s2 = std::move(s2);
    t->w2 = std::move(w2);
    t->norms_valid = true;
    t->cn8_valid = false;
}

// An input must hold finite values (checked on its node norms).
void check_finite(mwg_tree t, const char *what) {
    ensure_norms(t);
    for (size_t i = 0; i < t->s2.size(); i++)
        if (!std::isfinite(t->s2[i] + t->w2[i])) throw std::invalid_argument(std::string(what) + " holds NaN or Inf");
}

// Device memory for a result of n doubles, freed on failure.
struct DeviceResult {
    DeviceCtx *c = nullptr;
    double *p = nullptr;
    ~DeviceResult() { dev_free(c, p); }
    double *release() {
        double *q = p;
        p = nullptr;
        return q;
    }
};

// Room for n nodes in t (one allocation instead of repeated growth).
void reserve_nodes(Tree &t, size_t n) {
    t.coefs.reserve(n * (size_t)t.mra->block());
    t.keys.reserve(n);
    t.parent.reserve(n);
    t.child0.reserve(n);
    t.index.reserve(n);
}

// Where a node's block is on the device: a level's output buffer and the
// position in it.
struct NodeLoc {
    int level = -1;
    long long pos = 0;
};

// One apply of a batch and its refinement state.
struct Job {
    mwg_tree_s *out = nullptr;
    mwg_operator_s *op = nullptr;
    const Tree *inp = nullptr;
    mwg_tree inp_h = nullptr;
    const double *inp_cn8 = nullptr;  // the input's component norms (set by validate_jobs)
    std::vector<mwg_tree> prec_h;
    std::vector<PrecTree> prec_trees;  // filled by validate_jobs
    double prec = 0.0;
    int max_iter = -1;
    bool abs_prec = false;
    std::vector<int> work;        // output nodes of the current level
    std::vector<double> thr_all;  // threshold per output node
    double s_norm = 0.0, w_norm = 0.0;
    int iter = 0;
    ApplyPlan plan;
    std::string err;
};

void validate_jobs(std::vector<Job> &jobs) {
    std::set<const void *> outs, ins;
    for (const Job &j : jobs) ins.insert(j.inp);
    for (Job &j : jobs) {
        if (!(j.prec >= 0.0) || !std::isfinite(j.prec))
            throw std::invalid_argument("apply: precision must be finite and >= 0");
        check_compatible(j.out->tree, j.op->op, *j.inp);
        if (j.out->tree.mra != jobs[0].out->tree.mra)
            throw std::invalid_argument("apply: all applies of a batch must share one MRA");
        if (!outs.insert(&j.out->tree).second) throw std::invalid_argument("apply: an output tree appears twice");
        if (ins.count(&j.out->tree)) throw std::invalid_argument("apply: a tree is both output and input");
        for (mwg_tree t : j.prec_h)
            if (t->tree.mra != j.op->op.mra) throw std::invalid_argument("apply: prec tree on another MRA");
    }
    for (Job &j : jobs) {
        check_finite(j.inp_h, "apply: input tree");
        j.inp_cn8 = ensure_cnorms(j.inp_h);
        j.prec_trees.clear();
        for (mwg_tree t : j.prec_h) {
            if (outs.count(&t->tree)) throw std::invalid_argument("apply: a prec tree is an output of the batch");
            ensure_norms(t);
            j.prec_trees.push_back(PrecTree{&t->tree, &t->s2, &t->w2});
        }
    }
}

// The input trees' blocks of a batch on the device, one copy per distinct
// tree: a resident tree's own, else a temporary upload freed with the batch.
struct InputBlocks {
    TempBlocks tmp;
    std::map<mwg_tree, const double *> dev;
    explicit InputBlocks(mwg_context ctx) : tmp{ctx->dev, {}} {}
    const double *of(mwg_context ctx, mwg_tree t) {
        auto it = dev.find(t);
        if (it != dev.end()) return it->second;
        return dev[t] = tmp.of(ctx, t);
    }
};

// The device pools of one plan's parts, one after another in one
// allocation (the device plan adopts it; freed here otherwise).
struct LevelPool {
    DeviceCtx *c;
    double *base = nullptr;
    std::vector<double *> parts;
    ~LevelPool() { dev_free(c, base); }
    double *release() {
        double *b = base;
        base = nullptr;
        return b;
    }
};

// Host plans of `active` (in parallel across applies when there are several;
// build_plan parallelises internally otherwise), their pools made on the
// device from the input trees' blocks into one allocation, then the exact
// counts and segments (the fine single-component ones only for a mode with
// an FP64 class beside a low-precision one).
void build_plans(mwg_context ctx, std::vector<Job *> &active, const std::vector<std::vector<int>> &nodes,
                 const std::vector<std::vector<double>> &thr, InputBlocks &inputs, LevelPool &pools, int mode,
                 double *ms3 = nullptr) {
    const int n = (int)active.size();
    const double t0 = now_ms();
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 1) if (n > 1)
    for (int i = 0; i < n; i++) {
        Job &j = *active[i];
        try {
            build_plan(j.plan, j.op->op, *j.inp, j.out->tree, nodes[i], thr[i], 64, 256, true, j.inp_cn8);
        } catch (const std::exception &e) {
            j.err = e.what();
        }
    }
    for (Job *j : active)
        if (!j->err.empty()) throw std::runtime_error(j->err);
    const double t1 = now_ms();
    const MRA &mra = *active[0]->out->tree.mra;
    {
        size_t total = 0;
        for (Job *j : active) total += (size_t)j->plan.pool_values;
        std::string err;
        dev_free(pools.c, pools.base);
        pools.base = nullptr;
        dev_check(dev_alloc(pools.c, total, &pools.base, err), err);
        pools.parts.assign(n, nullptr);
        std::vector<PoolFill> fills(n);
        size_t off = 0;
        for (int i = 0; i < n; i++) {
            Job &j = *active[i];
            pools.parts[i] = pools.base + off;
            off += (size_t)j.plan.pool_values;
            fills[i] = PoolFill{inputs.of(ctx, j.inp_h), &j.plan, pools.parts[i]};
        }
        dev_check(dev_pool_fill(ctx->dev, mra, fills, err), err);
    }
    const double t2 = now_ms();
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 1) if (n > 1)
    for (int i = 0; i < n; i++) {
        Job &j = *active[i];
        try {
            finish_plan_counts(j.plan, j.op->op, 256, mode != MWG_MODE_FP64, n);
        } catch (const std::exception &e) {
            j.err = e.what();
        }
    }
    for (Job *j : active)
        if (!j->err.empty()) throw std::runtime_error(j->err);
    if (ms3) {
        ms3[0] += t1 - t0;
        ms3[1] += t2 - t1;
        ms3[2] += now_ms() - t2;
    }
}

// Adaptive applies with mrcpp::apply semantics per job; every refinement
// level of every job still refining is computed in one shared device plan.
int apply_batch(mwg_context ctx, std::vector<Job> &jobs) {
    if (!ctx || !ctx->dev) return fail(MWG_ERR_ARG, "apply: null context");
    if (jobs.empty()) return MWG_OK;
    validate_jobs(jobs);
    const MRA &mra = *jobs[0].out->tree.mra;
    const int blk = mra.block();
    // The output grids grow as structure only; their coefficients are sized
    // once at the end and filled by the download.  A failure leaves each
    // output as the zero function on its roots.
    struct OutputGuard {
        std::vector<Job> &jobs;
        bool done = false;
        ~OutputGuard() {
            if (done) return;
            for (Job &j : jobs) {
                Tree &o = j.out->tree;
                o.structure_only = false;
                o.clear();
                o.make_roots();
                host_changed(j.out);
                j.out->apply_thr.clear();
            }
        }
    } guard{jobs};
    for (Job &j : jobs) {
        Tree &o = j.out->tree;
        o.structure_only = false;
        o.clear();
        host_changed(j.out);
        // room for an output grid up to twice the input's size without
        // regrowing the structure (larger grids still grow)
        reserve_nodes(o, std::max<size_t>(64, 2 * (size_t)j.inp->size()));
        o.structure_only = true;
        o.make_roots();
        j.work.clear();
        for (int i = 0; i < o.size(); i++) j.work.push_back(i);
        j.thr_all.assign(o.size(), 0.0);
    }
    InputBlocks inputs(ctx);
    double kms = 0.0, pms = 0.0, slots = 0.0, gen = 0.0, computed = 0.0, apps_total = 0.0;
    double t_dev = 0.0, t_split = 0.0;  // device plan create + run + download; host split checks
    double t_create = 0.0, t_run = 0.0;  // parts of t_dev: plan creation, run (both host wall time)
    double plan3[3] = {0.0, 0.0, 0.0};   // parts of the plan time: grids and bands, device pools, counts
    const double t_all = now_ms();
    int launches = 0;
    // Every level's output blocks stay on the device until the post
    // transforms; each node records where its block is.
    struct LevelBufs {
        DeviceCtx *c;
        std::vector<double *> bufs;
        std::vector<long long> count;
        ~LevelBufs() {
            for (double *p : bufs) dev_free(c, p);
        }
    } keep{ctx->dev, {}, {}};
    std::vector<std::vector<NodeLoc>> locs(jobs.size());
    while (true) {
        std::vector<Job *> active;
        for (Job &j : jobs)
            if (!j.work.empty()) active.push_back(&j);
        if (active.empty()) break;
        std::vector<std::vector<int>> nodes(active.size());
        std::vector<std::vector<double>> thr(active.size());
        for (size_t a = 0; a < active.size(); a++) {
            Job &j = *active[a];
            const int M = j.op->op.size();
            const double Nsq = j.s_norm + j.w_norm;
            nodes[a] = j.work;
            thr[a].resize(j.work.size());
            for (size_t q = 0; q < j.work.size(); q++)
                thr[a][q] = j.prec * prec_factor(j.prec_trees, j.out->tree.keys[j.work[q]]) * std::sqrt(Nsq / M);
        }
        const double t0 = now_ms();
        LevelPool pools{ctx->dev};
        build_plans(ctx, active, nodes, thr, inputs, pools, ctx->mode, plan3);
        pms += now_ms() - t0;
        std::vector<PlanPart> parts;
        size_t total = 0;
        for (size_t a = 0; a < active.size(); a++) {
            Job *j = active[a];
            parts.push_back(PlanPart{j->op, &j->op->flat, &j->plan, pools.parts[a]});
            slots += (double)j->plan.slot_off.size();
            gen += (double)j->plan.n_generated;
            total += j->work.size();
        }
        std::string err;
        DevicePlan *dp = nullptr;
        const double t1 = now_ms();
        int rc = dev_plan_create(ctx->dev, parts, &dp, err, pools.base);
        if (rc) return dev_fail(rc, err);
        pools.release();  // the plan owns the pool now
        const double t1r = now_ms();
        t_create += t1r - t1;
        double ms = 0.0;
        unsigned long long apps = 0;
        rc = dev_plan_run(dp, ctx->mode, ctx->tau, &ms, &apps, nullptr, err, nullptr);
        t_run += now_ms() - t1r;
        // the level's blocks stay on the device; the host takes their norms
        double *lvl = rc ? nullptr : dev_plan_take_output(dp);
        dev_plan_destroy(dp);
        if (rc) return dev_fail(rc, err);
        const int level = (int)keep.bufs.size();
        keep.bufs.push_back(lvl);
        keep.count.push_back((long long)total);
        std::vector<double> s2(total), w2(total);
        rc = dev_node_norms(ctx->dev, mra.kp1(), lvl, (long long)total, s2.data(), w2.data(), err);
        if (rc) return dev_fail(rc, err);
        t_dev += now_ms() - t1;
        const double t2 = now_ms();
        kms += ms;
        apps_total += (double)apps;
        launches++;
        computed += (double)total;
        // Norms and split decisions: in parallel over the applies.  The
        // per-node squared norms are summed serially in node order, so the
        // thresholds and decisions do not depend on the thread count.
        std::vector<size_t> offs(active.size());
        for (size_t a = 0, off = 0; a < active.size(); a++) {
            offs[a] = off;
            off += active[a]->work.size();
        }
        const int na = (int)active.size();
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 1) if (na > 1)
        for (int a = 0; a < na; a++) {
            Job &j = *active[a];
            try {
                // j.plan stays: the next level's build reuses its pool storage
                Tree &o = j.out->tree;
                std::vector<NodeLoc> &loc = locs[&j - jobs.data()];
                const int max_depth = std::min(mra.max_depth, j.op->op.max_depth);
                const int nw = (int)j.work.size();
                loc.resize(o.size());
                for (int q = 0; q < nw; q++) {
                    const int i = j.work[q];
                    loc[i] = NodeLoc{level, (long long)(offs[a] + q)};
                    j.thr_all[i] = thr[a][q];
                    j.w_norm += w2[offs[a] + q];
                    if (o.keys[i].depth == 0) j.s_norm += s2[offs[a] + q];
                }
                const double tnorm = j.abs_prec ? 1.0 : std::sqrt(j.s_norm + j.w_norm);
                std::vector<int> splits;
                if (j.max_iter < 0 || j.iter < j.max_iter) {
                    for (int q = 0; q < nw; q++) {
                        const int i = j.work[q];
                        const NodeKey &k = o.keys[i];
                        if (k.depth >= max_depth) continue;
                        const double pf = prec_factor(j.prec_trees, k);
                        if (split_check(mra, k.depth, std::sqrt(w2[offs[a] + q]), j.prec * pf, tnorm))
                            splits.push_back(i);
                    }
                }
                const size_t grown = (size_t)o.size() + 8 * splits.size();
                o.keys.reserve(grown);
                o.parent.reserve(grown);
                o.child0.reserve(grown);
                std::vector<int> next;
                next.reserve(8 * splits.size());
                for (int i : splits) {
                    const int c0 = o.split(i);
                    for (int c = 0; c < 8; c++) next.push_back(c0 + c);
                }
                j.work.swap(next);
                j.thr_all.resize(o.size(), 0.0);
                j.iter++;
            } catch (const std::exception &e) {
                j.err = e.what();
            }
        }
        for (Job *j : active)
            if (!j->err.empty()) throw std::runtime_error(j->err);
        t_split += now_ms() - t2;
    }
    const double t3 = now_ms();
    {
        // Post transforms on the device over all levels' blocks: the 8
        // children of a branch are consecutive in the tree and were computed
        // consecutively in one level, so their blocks are adjacent.
        std::string err;
        int maxd = 0;
        for (Job &j : jobs) maxd = std::max(maxd, j.out->tree.max_depth());
        std::vector<std::vector<PostPair>> levels(maxd + 1);
        auto addr = [&](const NodeLoc &l) { return keep.bufs[l.level] + (size_t)l.pos * blk; };
        for (size_t ji = 0; ji < jobs.size(); ji++) {
            const Tree &o = jobs[ji].out->tree;
            const std::vector<NodeLoc> &loc = locs[ji];
            for (int i = 0; i < o.size(); i++) {
                const int c0 = o.child0[i];
                if (c0 < 0) continue;
                for (int c = 1; c < 8; c++)
                    if (loc[c0 + c].level != loc[c0].level || loc[c0 + c].pos != loc[c0].pos + c)
                        throw std::runtime_error("apply: a node's children were not computed together");
                levels[o.keys[i].depth].push_back(PostPair{addr(loc[i]), addr(loc[c0])});
            }
        }
        int rc = dev_post_transform(ctx->dev, mra, levels, err);
        // norms of the transformed blocks, level by level
        std::vector<std::vector<double>> s2(keep.bufs.size()), w2(keep.bufs.size());
        for (size_t l = 0; l < keep.bufs.size() && !rc; l++) {
            s2[l].resize(keep.count[l]);
            w2[l].resize(keep.count[l]);
            rc = dev_node_norms(ctx->dev, mra.kp1(), keep.bufs[l], keep.count[l], s2[l].data(), w2[l].data(), err);
        }
        // the blocks into the trees, in runs of nodes adjacent in both: into
        // the device copy of an output resident on this context, else
        // downloaded into the host tree
        for (size_t ji = 0; ji < jobs.size() && !rc; ji++) {
            mwg_tree out = jobs[ji].out;
            Tree &o = out->tree;
            const std::vector<NodeLoc> &loc = locs[ji];
            const bool resident = out->dctx == ctx;
            DeviceResult dres{ctx->dev, nullptr};
            if (resident) {
                rc = dev_alloc(ctx->dev, (size_t)o.size() * blk, &dres.p, err);
            } else {
                o.structure_only = false;
                o.coefs.resize((size_t)o.size() * blk);  // every block is written below
            }
            int i = 0;
            while (i < o.size() && !rc) {
                int e = i + 1;
                while (e < o.size() && loc[e].level == loc[i].level && loc[e].pos == loc[i].pos + (e - i)) e++;
                const size_t nv = (size_t)(e - i) * blk;
                if (resident) {
                    rc = dev_d2d(ctx->dev, dres.p + (size_t)i * blk, addr(loc[i]), nv, err);
                } else {
                    rc = dev_download(ctx->dev, addr(loc[i]), nv, o.node(i), err);
                    ctx->d2h_bytes += (long long)(sizeof(double) * nv);
                }
                i = e;
            }
            if (rc) break;
            std::vector<double> ns2(o.size()), nw2(o.size());
            for (int n = 0; n < o.size(); n++) {
                const NodeLoc &l = loc[n];
                ns2[n] = s2[l.level][l.pos];
                nw2[n] = w2[l.level][l.pos];
            }
            if (resident) {
                take_device_result(ctx, out, dres.release(), std::move(ns2), std::move(nw2));
            } else {
                double sq = 0.0;
                for (int n = 0; n < o.size(); n++) sq += nw2[n] + (o.parent[n] < 0 ? ns2[n] : 0.0);
                o.sqnorm = sq;
                host_changed(out);
                out->s2 = std::move(ns2);
                out->w2 = std::move(nw2);
                out->norms_valid = true;
            }
            out->apply_thr = jobs[ji].thr_all;
        }
        if (rc) return dev_fail(rc, err);
    }
    guard.done = true;
    auto &st = ctx->last_stats;
    st.assign(16, 0.0);
    st[0] = kms;
    st[1] = computed;
    st[2] = slots;
    st[3] = gen;
    st[4] = pms;
    st[5] = launches;
    st[6] = apps_total;
    st[7] = t_dev;
    st[8] = t_split;
    st[9] = now_ms() - t3;
    st[10] = now_ms() - t_all;
    st[11] = t_create;
    st[12] = t_run;
    st[13] = plan3[0];
    st[14] = plan3[1];
    st[15] = plan3[2];
    return MWG_OK;
}

// Frozen-grid device plan over the given applies (thr[i] per output node of
// out[i], tree order).
int make_plan(mwg_context ctx, int n, const mwg_tree *out, const mwg_operator *op, const mwg_tree *inp,
              const double *const *thr, int mode, mwg_plan *plan) {
    if (!ctx || !ctx->dev || n < 1 || !out || !op || !inp || !thr || !plan) return fail(MWG_ERR_ARG, "null argument");
    if (mode != MWG_MODE_FP64 && mode != MWG_MODE_MIXED && mode != MWG_MODE_SPLIT16)
        return fail(MWG_ERR_ARG, "unknown arithmetic mode");
    std::vector<Job> jobs(n);
    for (int i = 0; i < n; i++) {
        if (!out[i] || !op[i] || !inp[i] || !thr[i]) return fail(MWG_ERR_ARG, "null handle in batch");
        jobs[i].out = out[i];
        jobs[i].op = op[i];
        jobs[i].inp = &inp[i]->tree;
        jobs[i].inp_h = inp[i];
    }
    validate_jobs(jobs);
    auto p = std::make_unique<mwg_plan_s>();
    std::vector<Job *> active;
    std::vector<std::vector<int>> nodes(n);
    std::vector<std::vector<double>> th(n);
    for (int i = 0; i < n; i++) {
        const int no = out[i]->tree.size();
        nodes[i].resize(no);
        th[i].assign(thr[i], thr[i] + no);
        for (int g = 0; g < no; g++) {
            nodes[i][g] = g;
            if (!(th[i][g] >= 0.0) || !std::isfinite(th[i][g]))
                return fail(MWG_ERR_ARG, "thresholds must be finite and >= 0");
        }
        active.push_back(&jobs[i]);
    }
    const double t0 = now_ms();
    InputBlocks inputs(ctx);
    LevelPool pools{ctx->dev};
    build_plans(ctx, active, nodes, th, inputs, pools, mode);
    p->plan_ms = now_ms() - t0;
    std::vector<PlanPart> parts;
    for (int i = 0; i < n; i++) {
        const ApplyPlan &ap = jobs[i].plan;
        parts.push_back(PlanPart{op[i], &op[i]->flat, &ap, pools.parts[i]});
        p->slots += (double)ap.slot_off.size();
        p->gen += (double)ap.n_generated;
        p->host_apps += (double)ap.term_apps;
        p->segs += (double)ap.pseg_job.size();
        for (int w : ap.band_w) p->band_w = std::max(p->band_w, (double)w);
        p->outs.push_back(out[i]);
        p->ops.push_back(op[i]);
        p->keys.push_back(out[i]->tree.keys);
        p->n_out.push_back((int)nodes[i].size());
        p->total_out += (int)nodes[i].size();
    }
    p->thr = std::move(th);
    std::string err;
    int rc = dev_plan_create(ctx->dev, parts, &p->dev, err, pools.base);
    if (rc) return dev_fail(rc, err);
    pools.release();  // the plan owns the pool now
    {
        // (branch, first child) pairs of every output grid, by depth, as
        // indices into the plan's concatenated output nodes
        int maxd = 0;
        for (int i = 0; i < n; i++) maxd = std::max(maxd, out[i]->tree.max_depth());
        std::vector<std::vector<std::pair<long long, long long>>> levels(maxd + 1);
        long long off = 0;
        for (int i = 0; i < n; i++) {
            const Tree &o = out[i]->tree;
            for (int g = 0; g < o.size(); g++)
                if (o.child0[g] >= 0) levels[o.keys[g].depth].emplace_back(off + g, off + o.child0[g]);
            off += o.size();
        }
        rc = dev_plan_set_post(p->dev, *out[0]->tree.mra, levels, err);
        if (rc) {
            dev_plan_destroy(p->dev);
            return dev_fail(rc, err);
        }
    }
    p->ctx = ctx;
    p->mode = mode;
    p->tau = ctx->tau;
    p->tune_key = std::make_tuple(op[0]->flat.kp1, op[0]->flat.M, (int)p->band_w, mode);
    auto it = ctx->tuned.find(p->tune_key);
    if (it != ctx->tuned.end()) {
        p->tau = it->second.tau;
        rc = dev_plan_set_knobs(p->dev, it->second.single_ratio, it->second.order, err);
        if (rc) {
            dev_plan_destroy(p->dev);
            return dev_fail(rc, err);
        }
    }
    p->hold();
    p->held = true;
    *plan = p.release();
    return MWG_OK;
}

bool same_keys(const std::vector<NodeKey> &a, const std::vector<NodeKey> &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (a[i].depth != b[i].depth || a[i].l[0] != b[i].l[0] || a[i].l[1] != b[i].l[1] || a[i].l[2] != b[i].l[2])
            return false;
    return true;
}

// A plan's result into its output trees, finished: the post transforms run
// on the device on a copy of the output blocks (the plan keeps its NS blocks
// for further executions and downloads).
int plan_finish(mwg_plan plan) {
    for (size_t i = 0; i < plan->outs.size(); i++)
        if (!same_keys(plan->outs[i]->tree.keys, plan->keys[i]))
            return fail(MWG_ERR_ARG, "output tree grid changed since the plan was made");
    const MRA &mra = *plan->outs[0]->tree.mra;
    const int blk = mra.block();
    std::string err;
    DeviceCtx *dc = plan->ctx->dev;
    // the results go into the host trees (a resident output's device copy
    // becomes stale)
    for (mwg_tree o : plan->outs) {
        o->tree.structure_only = false;
        o->tree.coefs.resize(tree_values(o));
        host_changed(o);
        plan->ctx->d2h_bytes += (long long)(sizeof(double) * tree_values(o));
    }
    if (plan->post_done) {
        // the blocks were transformed on the device by mwg_plan_apply
        std::vector<double> s2(plan->total_out), w2(plan->total_out);
        int rc = dev_plan_download_norms(plan->dev, s2.data(), w2.data(), err);
        long long base = 0;
        for (size_t i = 0; i < plan->outs.size() && !rc; i++) {
            Tree &o = plan->outs[i]->tree;
            rc = dev_plan_download_range(plan->dev, base, plan->n_out[i], o.coefs.data(), err);
            double sq = 0.0;
            for (int g = 0; g < o.size(); g++) {
                sq += w2[base + g];
                if (o.parent[g] < 0) sq += s2[base + g];
            }
            o.sqnorm = sq;
            plan->outs[i]->apply_thr = plan->thr[i];
            base += plan->n_out[i];
        }
        if (rc) return dev_fail(rc, err);
        return MWG_OK;
    }
    double *tmp = nullptr;
    int rc = dev_plan_clone_output(plan->dev, &tmp, err);
    if (rc) return dev_fail(rc, err);
    struct Free {
        DeviceCtx *c;
        double *p;
        ~Free() { dev_free(c, p); }
    } hold{dc, tmp};
    // output node g of apply i sits at (base_i + g) in the plan's order
    int maxd = 0;
    for (auto *o : plan->outs) maxd = std::max(maxd, o->tree.max_depth());
    std::vector<std::vector<PostPair>> levels(maxd + 1);
    std::vector<size_t> base(plan->outs.size());
    for (size_t i = 0, off = 0; i < plan->outs.size(); i++) {
        base[i] = off;
        const Tree &o = plan->outs[i]->tree;
        for (int g = 0; g < o.size(); g++)
            if (o.child0[g] >= 0)
                levels[o.keys[g].depth].push_back(
                    PostPair{tmp + (off + g) * (size_t)blk, tmp + (off + o.child0[g]) * (size_t)blk});
        off += (size_t)plan->n_out[i];
    }
    rc = dev_post_transform(dc, mra, levels, err);
    std::vector<double> s2(plan->total_out), w2(plan->total_out);
    if (!rc) rc = dev_node_norms(dc, mra.kp1(), tmp, plan->total_out, s2.data(), w2.data(), err);
    for (size_t i = 0; i < plan->outs.size() && !rc; i++) {
        Tree &o = plan->outs[i]->tree;
        rc = dev_download(dc, tmp + base[i] * blk, (size_t)plan->n_out[i] * blk, o.coefs.data(), err);
        double sq = 0.0;
        for (int g = 0; g < o.size(); g++) {
            sq += w2[base[i] + g];
            if (o.parent[g] < 0) sq += s2[base[i] + g];
        }
        o.sqnorm = sq;
        plan->outs[i]->apply_thr = plan->thr[i];
    }
    if (rc) return dev_fail(rc, err);
    return MWG_OK;
}

}  // namespace

extern "C" {

const char *mwg_last_error(void) { return g_err.c_str(); }

const char *mwg_context_last_error(mwg_context ctx) { return ctx ? ctx->last_err.c_str() : ""; }

int mwg_context_create(int device, mwg_context *ctx) {
    return guarded([&] {
        if (!ctx) return fail(MWG_ERR_ARG, "null output pointer");
        std::string err;
        DeviceCtx *d = nullptr;
        int rc = dev_create(device, &d, err);
        if (rc) return fail(MWG_ERR_DEVICE, err);
        auto *c = new mwg_context_s;
        c->dev = d;
        {
            std::lock_guard<std::mutex> g(g_ctx_mtx);
            g_contexts.insert(c);
        }
        *ctx = c;
        return MWG_OK;
    });
}

int mwg_context_destroy(mwg_context ctx) {
    if (!ctx) return MWG_OK;
    if (ctx->plan_refs > 0) return fail(MWG_ERR_ARG, "context destroy: a plan on this context is still alive");
    if (ctx->resident_refs > 0)
        return fail(MWG_ERR_ARG, "context destroy: trees are still resident on this context (mwg_tree_set_resident)");
    {
        std::lock_guard<std::mutex> g(g_ctx_mtx);
        g_contexts.erase(ctx);
    }
    dev_destroy(ctx->dev);
    delete ctx;
    return MWG_OK;
}

int mwg_mra_create(int order, int basis_type, int root_scale, const int *corner, const int *boxes,
                   double scaling_factor, int max_depth, mwg_mra *out) {
    return guarded([&] {
        if (!out || !corner || !boxes) return fail(MWG_ERR_ARG, "null pointer argument");
        if (!(scaling_factor > 0.0)) return fail(MWG_ERR_ARG, "scaling factor must be > 0");
        if (max_depth < 1 || max_depth > 20) return fail(MWG_ERR_ARG, "max_depth must be in [1, 20]");
        for (int d = 0; d < 3; d++)
            if (boxes[d] < 1 || boxes[d] > 2) return fail(MWG_ERR_ARG, "boxes per direction must be 1 or 2");
        auto *m = new mwg_mra_s;
        try {
            m->mra.basis = make_basis(order, basis_type);
        } catch (...) {
            delete m;
            throw;
        }
        m->mra.n0 = root_scale;
        for (int d = 0; d < 3; d++) {
            m->mra.corner[d] = corner[d];
            m->mra.boxes[d] = boxes[d];
        }
        m->mra.sfac = scaling_factor;
        m->mra.max_depth = max_depth;
        *out = m;
        return MWG_OK;
    });
}

int mwg_mra_destroy(mwg_mra mra) {
    delete mra;
    return MWG_OK;
}

int mwg_mra_filter(mwg_mra mra, double *F) {
    if (!mra || !F) return fail(MWG_ERR_ARG, "null pointer argument");
    std::memcpy(F, mra->mra.basis.F.data(), sizeof(double) * mra->mra.basis.F.size());
    return MWG_OK;
}

static int make_op(mwg_mra mra, double mu, double prec, mwg_operator *out) {
    return guarded([&] {
        if (!mra || !out) return fail(MWG_ERR_ARG, "null pointer argument");
        if (!(prec > 0.0 && prec < 1.0)) return fail(MWG_ERR_ARG, "build precision must be in (0, 1)");
        if (!(mu >= 0.0) || !std::isfinite(mu)) return fail(MWG_ERR_ARG, "mu must be finite and >= 0");
        auto *o = new mwg_operator_s;
        try {
            build_operator(o->op, mra->mra, mu, prec, mra->mra.max_depth);
            flatten_operator(o->op, o->flat);
        } catch (...) {
            delete o;
            throw;
        }
        *out = o;
        return MWG_OK;
    });
}

int mwg_poisson_create(mwg_mra mra, double build_prec, mwg_operator *op) { return make_op(mra, 0.0, build_prec, op); }

int mwg_helmholtz_create(mwg_mra mra, double mu, double build_prec, mwg_operator *op) {
    if (!(mu > 0.0)) return fail(MWG_ERR_ARG, "Helmholtz mu must be > 0");
    return make_op(mra, mu, build_prec, op);
}

int mwg_gaussian_operator_create(mwg_mra mra, int n_terms, const double *alpha, const double *beta, int max_depth,
                                 mwg_operator *out) {
    return guarded([&] {
        if (!mra || !out || !alpha || !beta || n_terms < 1) return fail(MWG_ERR_ARG, "bad arguments");
        std::vector<GaussTerm> g(n_terms);
        for (int i = 0; i < n_terms; i++) {
            if (!(alpha[i] > 0.0) || !std::isfinite(beta[i])) return fail(MWG_ERR_ARG, "terms need alpha > 0");
            g[i] = {alpha[i], beta[i]};
        }
        auto *o = new mwg_operator_s;
        try {
            build_operator_terms(o->op, mra->mra, g, max_depth < 0 ? mra->mra.max_depth : max_depth);
            o->op.build_prec = 1e-14;
            flatten_operator(o->op, o->flat);
        } catch (...) {
            delete o;
            throw;
        }
        *out = o;
        return MWG_OK;
    });
}

int mwg_operators_create_device(mwg_context ctx, mwg_mra mra, int n, const double *mu, double build_prec,
                                mwg_operator *ops) {
    return guarded_ctx(ctx, [&] {
        if (!ctx || !mra || !mu || !ops) return fail(MWG_ERR_ARG, "null pointer argument");
        if (n < 1) return fail(MWG_ERR_ARG, "need at least one operator");
        if (!(build_prec > 0.0 && build_prec < 1.0)) return fail(MWG_ERR_ARG, "build precision must be in (0, 1)");
        for (int i = 0; i < n; i++)
            if (!(mu[i] >= 0.0) || !std::isfinite(mu[i])) return fail(MWG_ERR_ARG, "mu must be finite and >= 0");
        std::vector<std::unique_ptr<mwg_operator_s>> made(n);
        for (int i = 0; i < n; i++) made[i].reset(new mwg_operator_s);
        // kernel fits, one per operator, in parallel
        std::vector<std::string> fit_err(n);
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 1)
        for (int i = 0; i < n; i++) {
            try {
                prepare_operator(made[i]->op, mra->mra, mu[i], build_prec, mra->mra.max_depth);
            } catch (const std::exception &e) {
                fit_err[i] = e.what();
            }
        }
        for (int i = 0; i < n; i++)
            if (!fit_err[i].empty()) return fail(MWG_ERR_ARG, fit_err[i]);
        std::vector<OpBuildJob> jobs;
        for (int i = 0; i < n; i++) {
            made[i]->dev_ctx = ctx;
            jobs.push_back(OpBuildJob{&made[i]->op, &made[i]->flat, made[i].get()});
        }
        std::string err;
        int rc = dev_build_operators(ctx->dev, mra->mra, jobs, err);
        if (rc) {
            for (auto &o : made) dev_forget_operator(ctx->dev, o.get());
            return dev_fail(rc, err);
        }
        for (int i = 0; i < n; i++) ops[i] = made[i].release();
        return MWG_OK;
    });
}

namespace {
// Source of output node `k` in input tree f: its deepest node on the path
// from the roots, the levels below it and the child index of each level.
bool mul_source(const Tree &f, const NodeKey &k, long long &off, int &dep, unsigned &path) {
    for (int d = k.depth; d >= 0; d--) {
        long long l[3];
        for (int q = 0; q < 3; q++) l[q] = k.l[q] >> (k.depth - d);
        const int idx = f.find(d, l);
        if (idx < 0) continue;
        dep = k.depth - d;
        if (dep > 10) return false;
        path = 0;
        for (int lv = 1; lv <= dep; lv++) {
            unsigned t = 0;
            for (int q = 0; q < 3; q++) t |= (unsigned)((k.l[q] >> (k.depth - d - lv)) & 1) << q;
            path |= t << (3 * (lv - 1));
        }
        off = (long long)idx * f.mra->block();
        return true;
    }
    return false;
}

// Source node of input f for the sum at output node k: f's deepest node on
// the path, its block offset (base + index x block), the levels below it and
// the child index of each level.  An input without any node on the path
// (an empty tree) gets off = -1.  False when the node lies more than
// kAddMaxLevels levels below f's finest node on the path.
bool add_source(const Tree &f, const NodeKey &k, long long base, AddSrc &sr) {
    sr.off = -1;
    sr.dep = 0;
    sr.pad_ = 0;
    sr.path = 0;
    for (int d = k.depth; d >= 0; d--) {
        long long l[3];
        for (int q = 0; q < 3; q++) l[q] = k.l[q] >> (k.depth - d);
        const int idx = f.find(d, l);
        if (idx < 0) continue;
        sr.dep = k.depth - d;
        if (sr.dep > kAddMaxLevels) return false;
        for (int lv = 1; lv <= sr.dep; lv++) {
            unsigned long long t = 0;
            for (int q = 0; q < 3; q++) t |= (unsigned long long)((k.l[q] >> (k.depth - d - lv)) & 1) << q;
            sr.path |= t << (3 * (lv - 1));
        }
        sr.off = base + (long long)idx * f.mra->block();
        return true;
    }
    return true;
}

// Nodes of f in order of depth, each after its parent.
std::vector<int> by_depth(const Tree &f) {
    std::vector<int> order(f.size());
    for (int i = 0; i < f.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return f.keys[x].depth < f.keys[y].depth; });
    return order;
}

// Blocks of a tree grown level by level on the device (products and sums
// whose output stays on the device): each level's blocks in a buffer of its
// own, their norms on the host for the split decisions, and at the end all
// blocks in one array in tree order.
struct DeviceLevels {
    DeviceCtx *c;
    int blk, kp1;
    std::vector<double *> bufs;
    std::vector<std::vector<int>> nodes;  // tree nodes of each level, in buffer order
    std::vector<double> s2, w2;           // per tree node, of its last computed block
    ~DeviceLevels() {
        for (double *p : bufs) dev_free(c, p);
    }
    double *add_level(const std::vector<int> &work) {
        double *p = nullptr;
        std::string err;
        dev_check(dev_alloc(c, work.size() * (size_t)blk, &p, err), err);
        bufs.push_back(p);
        nodes.push_back(work);
        return p;
    }
    // the norms of the last level's blocks, tree_size nodes in the tree now
    void level_norms(int tree_size) {
        s2.resize(tree_size, 0.0);
        w2.resize(tree_size, 0.0);
        const std::vector<int> &w = nodes.back();
        std::vector<double> a(w.size()), b(w.size());
        std::string err;
        dev_check(dev_node_norms(c, kp1, bufs.back(), (long long)w.size(), a.data(), b.data(), err), err);
        for (size_t q = 0; q < w.size(); q++) {
            s2[w[q]] = a[q];
            w2[w[q]] = b[q];
        }
    }
    // every level's blocks into one array of tree_size blocks in tree order
    double *assemble(int tree_size) {
        DeviceResult r{c, nullptr};
        std::string err;
        dev_check(dev_alloc(c, (size_t)tree_size * blk, &r.p, err), err);
        for (size_t l = 0; l < bufs.size(); l++) {
            dev_check(dev_scatter_blocks(c, bufs[l], nodes[l], blk, r.p, err), err);
            dev_free(c, bufs[l]);
            bufs[l] = nullptr;
        }
        return r.release();
    }
};

// A failed product or sum leaves its output as the zero function on its
// roots.
struct ResetOnFailure {
    mwg_tree t;
    bool done = false;
    ~ResetOnFailure() {
        if (done) return;
        t->tree.structure_only = false;
        t->tree.clear();
        t->tree.make_roots();
        host_changed(t);
    }
};

// An input tree's blocks for a product or sum on ctx: in device memory when
// the tree is resident on ctx (on_dev), else the host copy.  noalloc: never
// an upload into a resident copy (that could allocate).
const double *input_blocks(mwg_context ctx, mwg_tree t, TempBlocks &tb, bool noalloc, bool &on_dev) {
    if (t->dctx == ctx && (t->dev_valid || !noalloc)) {
        on_dev = true;
        return noalloc ? t->dblocks : tb.of(ctx, t);
    }
    ensure_host(t);
    on_dev = false;
    ctx->h2d_bytes += (long long)(sizeof(double) * tree_values(t));
    return t->tree.coefs.data();
}

// The union of the grids of ins in T (which holds its roots).
int union_grid(Tree &T, const std::vector<const Tree *> &ins) {
    for (const Tree *f : ins) {
        std::vector<int> order(f->size());
        for (int i = 0; i < f->size(); i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return f->keys[x].depth < f->keys[y].depth; });
        for (int i : order) {
            const NodeKey &k = f->keys[i];
            if (k.depth == 0 || T.find(k.depth, k.l) >= 0) continue;
            long long lp[3] = {k.l[0] >> 1, k.l[1] >> 1, k.l[2] >> 1};
            const int par = T.find(k.depth - 1, lp);
            if (par < 0) return fail(MWG_ERR_ARG, "input tree is not complete from its roots");
            T.split(par);
        }
    }
    return MWG_OK;
}

// T = sum_i coefs[i] ins[i] (mrcpp::add).  prec < 0: on the union of the
// input grids, every node computed in one launch.  prec >= 0: built from the
// roots, level by level, splitting a node when the sum's wavelet norm fails
// the split check (absolute, or relative to the sum's norm estimated from the
// leaves), for at most max_iter levels (< 0: no limit) and the MRA's depth.
// With dres the sum is built on the device and its blocks (tree order) and
// node norms are returned there; otherwise each level's blocks are
// downloaded into T.
struct DeviceSum {
    double *blocks = nullptr;
    std::vector<double> s2, w2;
};
int add_trees(mwg_context ctx, double prec, Tree &T, const std::vector<mwg_tree> &ins,
              const std::vector<double> &coefs, int max_iter, bool abs_prec, DeviceSum *dres) {
    const MRA &mra = *T.mra;
    const int blk = mra.block(), n3 = mra.kp1_3();
    const bool on_device = dres != nullptr;
    TempBlocks tb{ctx->dev, {}};
    std::vector<BlockSource> pool;
    std::vector<long long> base(ins.size());
    size_t total = 0;
    for (size_t i = 0; i < ins.size(); i++) {
        base[i] = (long long)total;
        bool on_dev = false;
        const double *p = input_blocks(ctx, ins[i], tb, false, on_dev);
        pool.push_back(BlockSource{p, tree_values(ins[i]), on_dev});
        total += tree_values(ins[i]);
    }
    T.structure_only = false;
    T.clear();
    // the union grid holds at most all input nodes
    if (prec < 0.0) reserve_nodes(T, total / (size_t)blk);
    T.structure_only = on_device;
    T.make_roots();
    if (prec < 0.0) {
        std::vector<const Tree *> grids;
        for (mwg_tree f : ins) grids.push_back(&f->tree);
        if (int rc = union_grid(T, grids)) return rc;
    }
    std::vector<int> work;
    if (prec < 0.0) {
        for (int i = 0; i < T.size(); i++) work.push_back(i);
    } else {
        for (int i = 0; i < T.size(); i++)
            if (T.keys[i].depth == 0) work.push_back(i);
    }
    std::string err;
    DevAddSession *m = nullptr;
    dev_check(dev_add_begin(ctx->dev, mra, pool, coefs, &m, err), err);
    std::unique_ptr<DevAddSession, void (*)(DevAddSession *)> session(m, dev_add_end);
    DeviceLevels lv{ctx->dev, blk, mra.kp1(), {}, {}, {}, {}};
    const size_t n_in = ins.size();
    for (int iter = 0;; iter++) {
        std::vector<AddSrc> src(work.size() * n_in);
        for (size_t w = 0; w < work.size(); w++)
            for (size_t i = 0; i < n_in; i++)
                if (!add_source(ins[i]->tree, T.keys[work[w]], base[i], src[w * n_in + i]))
                    return fail(MWG_ERR_ARG, "output node more than " + std::to_string(kAddMaxLevels) +
                                                 " levels below an input's finest node");
        if (on_device) {
            double *buf = lv.add_level(work);
            dev_check(dev_add_nodes(m, src, buf, err, true), err);
            lv.level_norms(T.size());
        } else {
            std::vector<double> blocks(work.size() * (size_t)blk);
            dev_check(dev_add_nodes(m, src, blocks.data(), err), err);
            ctx->d2h_bytes += (long long)(sizeof(double) * blocks.size());
#pragma omp parallel for num_threads(team_size()) schedule(static)
            for (long long w = 0; w < (long long)work.size(); w++)
                std::copy(blocks.begin() + w * blk, blocks.begin() + (w + 1) * blk, T.node(work[w]));
        }
        if (prec < 0.0 || (max_iter >= 0 && iter >= max_iter)) break;
        double tnorm = 1.0;
        if (!abs_prec) {
            double sq = 0.0;
            for (int i = 0; i < T.size(); i++)
                if (T.child0[i] < 0) {
                    if (on_device) {
                        sq += lv.s2[i];
                    } else {
                        for (int e = 0; e < n3; e++) sq += T.node(i)[e] * T.node(i)[e];
                    }
                }
            tnorm = std::sqrt(sq);
        }
        std::vector<int> next;
        for (int wi : work) {
            const int d = T.keys[wi].depth;
            double wsq = 0.0;
            if (on_device) {
                wsq = lv.w2[wi];
            } else {
                for (int e = n3; e < blk; e++) wsq += T.node(wi)[e] * T.node(wi)[e];
            }
            if (d < mra.max_depth && split_check(mra, d, std::sqrt(wsq), prec, tnorm)) {
                const int c0 = T.split(wi);
                for (int t = 0; t < 8; t++) next.push_back(c0 + t);
            }
        }
        if (next.empty()) break;
        work.swap(next);
    }
    session.reset();
    if (on_device) {
        // every node was computed once: its level's norms are its own
        dres->blocks = lv.assemble(T.size());
        dres->s2 = std::move(lv.s2);
        dres->w2 = std::move(lv.w2);
        dres->s2.resize(T.size(), 0.0);
        dres->w2.resize(T.size(), 0.0);
    } else {
        T.sqnorm = tree_square_norm(T);
    }
    return MWG_OK;
}
}  // namespace

namespace {
// Max norms of a tree for the product's max-norm refinement: per node, the
// largest 2^{3n} ||block||^2 (ms) and 2^{3n} ||wavelet part||^2 (mw) over the
// node and its descendants, n = the node's scale (root scale + depth): the
// scaled squared norm estimates the squared sup-norm over the node's box.
void max_norms(mwg_tree t, std::vector<double> &ms, std::vector<double> &mw) {
    ensure_norms(t);
    const Tree &T = t->tree;
    const int n = T.size();
    ms.assign(n, 0.0);
    mw.assign(n, 0.0);
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return T.keys[x].depth > T.keys[y].depth; });
    for (int i : order) {
        const double f = std::pow(2.0, 3.0 * (T.mra->n0 + T.keys[i].depth));
        ms[i] = std::max(ms[i], f * (t->s2[i] + t->w2[i]));
        mw[i] = std::max(mw[i], f * t->w2[i]);
        if (T.parent[i] >= 0) {
            ms[T.parent[i]] = std::max(ms[T.parent[i]], ms[i]);
            mw[T.parent[i]] = std::max(mw[T.parent[i]], mw[i]);
        }
    }
}

// The max norms of input f at output node k: f's node at k, or below f's
// finest node on the path that node's scaling max with no wavelet part
// (f is a polynomial there).
void max_norms_at(const Tree &f, const std::vector<double> &ms, const std::vector<double> &mw, const NodeKey &k,
                  double &s, double &w) {
    s = w = 0.0;
    for (int d = k.depth; d >= 0; d--) {
        long long l[3];
        for (int q = 0; q < 3; q++) l[q] = k.l[q] >> (k.depth - d);
        const int idx = f.find(d, l);
        if (idx < 0) continue;
        s = std::sqrt(ms[idx]);
        w = d == k.depth ? std::sqrt(mw[idx]) : 0.0;
        return;
    }
}

// mrcpp::multiply on the device: the output grid from the union of the
// inputs' grids, refined level by level with the split rule.  With
// max_norms (useMaxNorms) the grid instead grows from the roots with the
// product rule of the reference's MultiplicationAdaptor: a node splits when
// maxW_a maxS_b + maxW_b maxS_a > prec, from the inputs' max norms.  noalloc: the
// device buffers come from the device's scratch array.  An output resident
// on ctx is built on the device (levels, BottomUp transform and norms there)
// and only the node norms come to the host; otherwise each level's blocks are
// downloaded and the tree is compressed on the host.
int multiply_trees(mwg_context ctx, double prec, mwg_tree out, double coef, mwg_tree a, mwg_tree b, int max_iter,
                   int abs_prec, bool noalloc, bool use_max_norms = false) {
    return guarded_ctx(ctx, [&] {
        if (!ctx || !out || !a || !b) return fail(MWG_ERR_ARG, "null pointer argument");
        if (out == a || out == b) return fail(MWG_ERR_ARG, "the output tree must differ from the inputs");
        const MRA &mra = *out->tree.mra;
        if (a->tree.mra != &mra || b->tree.mra != &mra) return fail(MWG_ERR_ARG, "trees of different MRAs");
        if (mra.basis.type != BASIS_INTERPOLATING) return fail(MWG_ERR_ARG, "multiplication needs the interpolating basis");
        if (!(prec >= 0.0) || !std::isfinite(coef)) return fail(MWG_ERR_ARG, "bad precision or coefficient");
        const bool on_device = out->dctx == ctx && !noalloc;
        Tree &T = out->tree;
        T.structure_only = false;
        T.clear();
        host_changed(out);
        ResetOnFailure reset{out};
        // the union grid holds at most all input nodes
        reserve_nodes(T, (size_t)a->tree.size() + (size_t)b->tree.size());
        T.structure_only = on_device;
        T.make_roots();
        std::vector<double> ms_a, mw_a, ms_b, mw_b;
        if (use_max_norms) {
            max_norms(a, ms_a, mw_a);
            max_norms(b, ms_b, mw_b);
        } else if (int rc = union_grid(T, {&a->tree, &b->tree})) {
            return rc;
        }
        std::vector<int> work;
        for (int i = 0; i < T.size(); i++)
            if (T.child0[i] < 0) work.push_back(i);
        std::string err;
        TempBlocks tb{ctx->dev, {}};
        bool a_dev = false, b_dev = false;
        const double *pa = input_blocks(ctx, a, tb, noalloc, a_dev);
        const double *pb = input_blocks(ctx, b, tb, noalloc, b_dev);
        DevMulSession *m = nullptr;
        dev_check(dev_mul_begin(ctx->dev, mra, pa, tree_values(a), pb, tree_values(b), &m, err, noalloc, a_dev, b_dev),
                  err);
        std::unique_ptr<DevMulSession, void (*)(DevMulSession *)> session(m, dev_mul_end);
        const int blk = mra.block(), n3 = mra.kp1_3();
        DeviceLevels lv{ctx->dev, blk, mra.kp1(), {}, {}, {}, {}};
        for (int iter = 0;; iter++) {
            std::vector<MulNode> nodes(work.size());
            for (size_t w = 0; w < work.size(); w++) {
                const NodeKey &k = T.keys[work[w]];
                MulNode &nd = nodes[w];
                if (!mul_source(a->tree, k, nd.src_a, nd.dep_a, nd.path_a) ||
                    !mul_source(b->tree, k, nd.src_b, nd.dep_b, nd.path_b))
                    return fail(MWG_ERR_ARG, "output node more than 10 levels below an input's finest node");
                nd.hfac = std::pow(mra.box_size(k.depth + 1), -1.5);
            }
            if (on_device) {
                double *buf = lv.add_level(work);
                dev_check(dev_mul_nodes(m, nodes, coef, buf, err, true), err);
                lv.level_norms(T.size());
            } else {
                std::vector<double> blocks(work.size() * (size_t)blk);
                dev_check(dev_mul_nodes(m, nodes, coef, blocks.data(), err), err);
                ctx->d2h_bytes += (long long)(sizeof(double) * blocks.size());
#pragma omp parallel for num_threads(team_size()) schedule(static)
                for (long long w = 0; w < (long long)work.size(); w++)
                    std::copy(blocks.begin() + w * blk, blocks.begin() + (w + 1) * blk, T.node(work[w]));
            }
            if (max_iter >= 0 && iter >= max_iter) break;
            if (use_max_norms) {
                std::vector<int> next;
                for (int wi : work) {
                    const NodeKey k = T.keys[wi];
                    if (k.depth >= mra.max_depth) continue;
                    double sa, wa, sb, wb;
                    max_norms_at(a->tree, ms_a, mw_a, k, sa, wa);
                    max_norms_at(b->tree, ms_b, mw_b, k, sb, wb);
                    if (wa * sb + wb * sa > prec) {
                        const int c0 = T.split(wi);
                        for (int t = 0; t < 8; t++) next.push_back(c0 + t);
                    }
                }
                if (next.empty()) break;
                work.swap(next);
                continue;
            }
            // norm estimate of the product from the leaves' scaling parts
            double tnorm = 1.0;
            if (!abs_prec) {
                double sq = 0.0;
                for (int i = 0; i < T.size(); i++)
                    if (T.child0[i] < 0) {
                        if (on_device) {
                            sq += lv.s2[i];
                        } else {
                            for (int e = 0; e < n3; e++) sq += T.node(i)[e] * T.node(i)[e];
                        }
                    }
                tnorm = std::sqrt(sq);
            }
            std::vector<int> next;
            for (int wi : work) {
                const int d = T.keys[wi].depth;
                double wsq = 0.0;
                if (on_device) {
                    wsq = lv.w2[wi];
                } else {
                    for (int e = n3; e < blk; e++) wsq += T.node(wi)[e] * T.node(wi)[e];
                }
                if (d < mra.max_depth && split_check(mra, d, std::sqrt(wsq), prec, tnorm)) {
                    const int c0 = T.split(wi);
                    for (int t = 0; t < 8; t++) next.push_back(c0 + t);
                }
            }
            if (next.empty()) break;
            work.swap(next);
        }
        session.reset();
        if (on_device) {
            DeviceResult res{ctx->dev, lv.assemble(T.size())};
            // BottomUp: every branch's block from its children's scaling parts
            std::vector<std::vector<PostPair>> levels(T.max_depth() + 1);
            for (int i = 0; i < T.size(); i++)
                if (T.child0[i] >= 0)
                    levels[T.keys[i].depth].push_back(
                        PostPair{res.p + (size_t)i * blk, res.p + (size_t)T.child0[i] * blk});
            dev_check(dev_post_transform(ctx->dev, mra, levels, err, false), err);
            std::vector<double> s2(T.size()), w2(T.size());
            dev_check(dev_node_norms(ctx->dev, mra.kp1(), res.p, T.size(), s2.data(), w2.data(), err), err);
            take_device_result(ctx, out, res.release(), std::move(s2), std::move(w2));
        } else {
            tree_compress(T);
            T.sqnorm = tree_square_norm(T);
        }
        reset.done = true;
        return (int)MWG_OK;
    });
}
}  // namespace

int mwg_multiply(mwg_context ctx, double prec, mwg_tree out, double coef, mwg_tree a, mwg_tree b, int max_iter,
                 int abs_prec) {
    return multiply_trees(ctx, prec, out, coef, a, b, max_iter, abs_prec, false);
}

int mwg_multiply_noalloc(mwg_context ctx, double prec, mwg_tree out, double coef, mwg_tree a, mwg_tree b,
                         int max_iter, int abs_prec) {
    return multiply_trees(ctx, prec, out, coef, a, b, max_iter, abs_prec, true);
}

int mwg_multiply_ex(mwg_context ctx, double prec, mwg_tree out, double coef, mwg_tree a, mwg_tree b, int max_iter,
                    int abs_prec, int flags) {
    if (flags & ~MWG_MUL_MAX_NORMS) return fail(MWG_ERR_ARG, "unknown multiply flags");
    return multiply_trees(ctx, prec, out, coef, a, b, max_iter, abs_prec, false, (flags & MWG_MUL_MAX_NORMS) != 0);
}

int mwg_add(mwg_context ctx, double prec, mwg_tree out, int n, const double *coefs, const mwg_tree *inputs,
            int max_iter, int abs_prec) {
    return guarded_ctx(ctx, [&] {
        if (!ctx || !out || !coefs || !inputs) return fail(MWG_ERR_ARG, "null pointer argument");
        if (n < 1) return fail(MWG_ERR_ARG, "at least one input tree is required");
        if (std::isnan(prec)) return fail(MWG_ERR_ARG, "bad precision");
        std::vector<mwg_tree> ins(n);
        std::vector<double> cs(coefs, coefs + n);
        for (int i = 0; i < n; i++) {
            if (!inputs[i]) return fail(MWG_ERR_ARG, "null input tree");
            if (inputs[i] == out) return fail(MWG_ERR_ARG, "the output tree must differ from the inputs");
            if (inputs[i]->tree.mra != out->tree.mra) return fail(MWG_ERR_ARG, "trees of different MRAs");
            if (!std::isfinite(cs[i])) return fail(MWG_ERR_ARG, "non-finite coefficient");
            ins[i] = inputs[i];
        }
        host_changed(out);
        ResetOnFailure reset{out};
        const bool on_device = out->dctx == ctx;
        DeviceSum ds;
        const int rc = add_trees(ctx, prec, out->tree, ins, cs, max_iter, abs_prec != 0, on_device ? &ds : nullptr);
        if (rc != MWG_OK) {
            dev_free(ctx->dev, ds.blocks);
            return rc;
        }
        if (on_device) take_device_result(ctx, out, ds.blocks, std::move(ds.s2), std::move(ds.w2));
        reset.done = true;
        return (int)MWG_OK;
    });
}

int mwg_tree_add(mwg_context ctx, mwg_tree tree, double coef, mwg_tree inp) {
    return guarded_ctx(ctx, [&] {
        if (!ctx || !tree || !inp) return fail(MWG_ERR_ARG, "null pointer argument");
        if (tree == inp) {
            // f + c f on f's own grid
            if (!std::isfinite(coef)) return fail(MWG_ERR_ARG, "non-finite coefficient");
            return mwg_tree_rescale(tree, 1.0 + coef);
        }
        if (inp->tree.mra != tree->tree.mra) return fail(MWG_ERR_ARG, "trees of different MRAs");
        if (!std::isfinite(coef)) return fail(MWG_ERR_ARG, "non-finite coefficient");
        Tree sum;
        sum.mra = tree->tree.mra;
        const bool on_device = tree->dctx == ctx;
        DeviceSum ds;
        const int rc = add_trees(ctx, -1.0, sum, {tree, inp}, {1.0, coef}, -1, true, on_device ? &ds : nullptr);
        if (rc != MWG_OK) {
            dev_free(ctx->dev, ds.blocks);
            return rc;
        }
        tree->tree = std::move(sum);
        if (on_device) {
            take_device_result(ctx, tree, ds.blocks, std::move(ds.s2), std::move(ds.w2));
        } else {
            host_changed(tree);
        }
        return (int)MWG_OK;
    });
}

int mwg_tree_crop(mwg_tree tree, double prec, double split_fac, int abs_prec) {
    return guarded([&] {
        if (!tree) return fail(MWG_ERR_ARG, "null tree");
        if (std::isnan(prec) || !(split_fac >= 0.0)) return fail(MWG_ERR_ARG, "bad precision or split factor");
        Tree &T = tree->tree;
        if (T.size() == 0 || !(prec > 0.0)) return (int)MWG_OK;
        const MRA &mra = *T.mra;
        const int blk = mra.block(), n3 = mra.kp1_3();
        // a tree whose blocks only the device holds is cropped there: the
        // decisions from its node norms, the kept blocks gathered on the
        // device
        const bool on_device = !tree->host_valid;
        if (on_device) ensure_norms(tree);
        auto sqnorm = [&]() {
            if (!on_device) return tree_square_norm(T);
            double sq = 0.0;
            for (int i = 0; i < T.size(); i++) sq += tree->w2[i] + (T.parent[i] < 0 ? tree->s2[i] : 0.0);
            return sq;
        };
        const double tnorm = abs_prec ? 1.0 : std::sqrt(std::max(sqnorm(), 0.0));
        // post-order: a branch whose children are all leaves (after their own
        // crop) and whose wavelet norm fails the split check loses them
        std::vector<char> keep(T.size(), 1), leaf(T.size(), 0);
        std::vector<int> order = by_depth(T);
        for (auto it = order.rbegin(); it != order.rend(); ++it) {
            const int i = *it;
            if (T.child0[i] < 0) {
                leaf[i] = 1;
                continue;
            }
            bool all_leaves = true;
            for (int t = 0; t < 8; t++) all_leaves = all_leaves && leaf[T.child0[i] + t];
            if (!all_leaves) continue;
            double wsq = 0.0;
            if (on_device) {
                wsq = tree->w2[i];
            } else {
                for (int e = n3; e < blk; e++) wsq += T.node(i)[e] * T.node(i)[e];
            }
            const int d = T.keys[i].depth;
            // the split check with the wavelet threshold's scale exponent
            // multiplied by split_fac (1: the WaveletAdaptor rule)
            const double scale = mra.n0 + d;
            const double thr = std::max(2.0 * 2.220446049250313e-16,
                                        prec * tnorm * std::pow(2.0, -0.5 * split_fac * (scale + 1)));
            if (!(std::sqrt(wsq) > thr)) {
                for (int t = 0; t < 8; t++) keep[T.child0[i] + t] = 0;
                leaf[i] = 1;
            }
        }
        // a removed node's descendants go with it
        for (int i : order)
            if (T.parent[i] >= 0 && !keep[T.parent[i]]) keep[i] = 0;
        Tree C;
        C.mra = T.mra;
        C.structure_only = on_device;
        C.make_roots();
        std::vector<int> src_of;  // node of T behind each node of C
        for (int i : order) {
            if (!keep[i]) continue;
            const NodeKey &k = T.keys[i];
            int j = C.find(k.depth, k.l);
            if (j < 0) {
                long long lp[3] = {k.l[0] >> 1, k.l[1] >> 1, k.l[2] >> 1};
                const int par = C.find(k.depth - 1, lp);
                if (par < 0) return fail(MWG_ERR_RUNTIME, "crop: parent of a kept node missing");
                C.split(par);
                j = C.find(k.depth, k.l);
            }
            if ((int)src_of.size() < C.size()) src_of.resize(C.size(), -1);
            src_of[j] = i;
            if (!on_device) std::copy(T.node(i), T.node(i) + blk, C.node(j));
        }
        if (!on_device) {
            C.sqnorm = tree_square_norm(C);
            T = std::move(C);
            host_changed(tree);
            return (int)MWG_OK;
        }
        for (int v : src_of)
            if (v < 0) return fail(MWG_ERR_RUNTIME, "crop: a node of the cropped tree has no source");
        mwg_context dc = tree->dctx;
        DeviceResult res{dc->dev, nullptr};
        std::string err;
        dev_check(dev_alloc(dc->dev, (size_t)C.size() * blk, &res.p, err), err);
        dev_check(dev_gather_blocks(dc->dev, tree->dblocks, src_of, blk, res.p, err), err);
        std::vector<double> s2(C.size()), w2(C.size());
        for (int j = 0; j < C.size(); j++) {
            s2[j] = tree->s2[src_of[j]];
            w2[j] = tree->w2[src_of[j]];
        }
        T = std::move(C);
        take_device_result(dc, tree, res.release(), std::move(s2), std::move(w2));
        return (int)MWG_OK;
    });
}

int mwg_operator_destroy(mwg_operator op) {
    if (!op) return MWG_OK;
    if (op->plan_refs > 0) return fail(MWG_ERR_ARG, "operator destroy: a plan still uses this operator");
    {
        std::lock_guard<std::mutex> g(g_ctx_mtx);
        for (auto *c : g_contexts) dev_forget_operator(c->dev, op);
    }
    delete op;
    return MWG_OK;
}

int mwg_operator_info(mwg_operator op, int *n_terms, int *n_depth, int *kp1, double *r_min, double *r_max,
                      double *fit_err) {
    if (!op) return fail(MWG_ERR_ARG, "null operator");
    if (n_terms) *n_terms = op->op.size();
    if (n_depth) *n_depth = op->op.max_depth + 1;
    if (kp1) *kp1 = op->op.mra->kp1();
    if (r_min) *r_min = op->op.r_min;
    if (r_max) *r_max = op->op.r_max;
    if (fit_err) *fit_err = op->op.fit_err;
    return MWG_OK;
}

int mwg_operator_terms(mwg_operator op, double *alpha, double *beta) {
    if (!op) return fail(MWG_ERR_ARG, "null operator");
    for (int m = 0; m < op->op.size(); m++) {
        if (alpha) alpha[m] = op->op.terms[m].alpha;
        if (beta) beta[m] = op->op.terms[m].beta;
    }
    return MWG_OK;
}

int mwg_operator_table_sizes(mwg_operator op, long long *nb, long long *nn) {
    if (!op) return fail(MWG_ERR_ARG, "null operator");
    if (nb) *nb = op->flat.device_blocks > 0 ? op->flat.device_blocks : (long long)op->flat.blocks.size();
    if (nn) *nn = (long long)op->flat.norms.size();
    return MWG_OK;
}

int mwg_operator_tables(mwg_operator op, int *L, long long *boff, long long *noff, double *blocks, double *norms) {
    if (!op) return fail(MWG_ERR_ARG, "null operator");
    const OperFlat &f = op->flat;
    if (L) std::memcpy(L, f.L.data(), sizeof(int) * f.L.size());
    if (boff) std::memcpy(boff, f.boff.data(), sizeof(long long) * f.boff.size());
    if (noff) std::memcpy(noff, f.noff.data(), sizeof(long long) * f.noff.size());
    if (blocks) {
        if (f.device_blocks > 0) {
            std::lock_guard<std::mutex> g(g_ctx_mtx);
            if (!g_contexts.count(op->dev_ctx))
                return fail(MWG_ERR_ARG, "the context holding this operator's tables was destroyed");
            std::string err;
            const int rc = dev_download_operator_blocks(op->dev_ctx->dev, op, blocks, f.device_blocks, err);
            if (rc) return dev_fail(rc, err);
        } else {
            std::memcpy(blocks, f.blocks.data(), sizeof(double) * f.blocks.size());
        }
    }
    if (norms) std::memcpy(norms, f.norms.data(), sizeof(double) * f.norms.size());
    return MWG_OK;
}

int mwg_tree_create(mwg_mra mra, mwg_tree *out) {
    if (!mra || !out) return fail(MWG_ERR_ARG, "null pointer argument");
    auto *t = new mwg_tree_s;
    t->tree.mra = &mra->mra;
    t->tree.make_roots();
    *out = t;
    return MWG_OK;
}

int mwg_tree_destroy(mwg_tree t) {
    if (t && t->plan_refs > 0) return fail(MWG_ERR_ARG, "tree destroy: a plan still writes into this tree");
    if (t && t->dctx) {
        dev_free(t->dctx->dev, t->dblocks);
        t->dctx->resident_refs--;
    }
    delete t;
    return MWG_OK;
}

int mwg_tree_set_resident(mwg_tree t, mwg_context ctx, int on) {
    return guarded_ctx(ctx, [&] {
        if (!t || (on && !ctx)) return fail(MWG_ERR_ARG, "null pointer argument");
        if (t->plan_refs > 0) return fail(MWG_ERR_ARG, "set_resident: a plan still writes into this tree");
        if (on && t->dctx == ctx) return (int)MWG_OK;
        if (t->dctx) {
            // leave the current device: the host copy first
            ensure_host(t);
            dev_free(t->dctx->dev, t->dblocks);
            t->dctx->resident_refs--;
            t->dctx = nullptr;
            t->dblocks = nullptr;
            t->dcap = 0;
            t->dev_valid = false;
        }
        if (!on) return (int)MWG_OK;
        // the blocks move to the device at the first call on ctx that reads
        // them (an output never uploads its old blocks)
        t->dctx = ctx;
        ctx->resident_refs++;
        return (int)MWG_OK;
    });
}

int mwg_tree_is_resident(mwg_tree t, int *on_host, int *on_device) {
    if (!t) return fail(MWG_ERR_ARG, "null tree");
    if (on_host) *on_host = t->host_valid ? 1 : 0;
    if (on_device) *on_device = t->dev_valid ? 1 : 0;
    return t->dctx ? 1 : 0;
}

int mwg_context_transfer_bytes(mwg_context ctx, long long *h2d, long long *d2h, int reset) {
    if (!ctx) return fail(MWG_ERR_ARG, "null context");
    if (h2d) *h2d = ctx->h2d_bytes.load();
    if (d2h) *d2h = ctx->d2h_bytes.load();
    if (reset) {
        ctx->h2d_bytes = 0;
        ctx->d2h_bytes = 0;
    }
    return MWG_OK;
}

int mwg_tree_project_gaussians(mwg_tree t, double prec, int abs_prec, int max_depth, int n, const double *coef,
                               const double *exponent, const double *center) {
    return guarded([&] {
        if (!t || n < 0 || (n > 0 && (!coef || !exponent || !center))) return fail(MWG_ERR_ARG, "bad arguments");
        GaussSum f;
        f.coef.assign(coef, coef + n);
        f.exp.assign(exponent, exponent + n);
        f.pos.assign(center, center + 3 * n);
        for (int i = 0; i < n; i++)
            if (!(exponent[i] > 0.0)) return fail(MWG_ERR_ARG, "Gaussian exponents must be > 0");
        t->tree.structure_only = false;
        project_adaptive(t->tree, f, prec, abs_prec != 0, max_depth);
        host_changed(t);
        t->apply_thr.clear();
        return MWG_OK;
    });
}

int mwg_tree_size(mwg_tree t, int *n) {
    if (!t || !n) return fail(MWG_ERR_ARG, "null pointer argument");
    *n = t->tree.size();
    return MWG_OK;
}

int mwg_tree_get(mwg_tree t, int *depth, long long *l, double *coefs) {
    return guarded([&] {
        if (!t) return fail(MWG_ERR_ARG, "null tree");
        const Tree &tr = t->tree;
        for (int i = 0; i < tr.size(); i++) {
            if (depth) depth[i] = tr.keys[i].depth;
            if (l)
                for (int d = 0; d < 3; d++) l[3 * i + d] = tr.keys[i].l[d];
        }
        if (coefs) {
            ensure_host(t);
            std::memcpy(coefs, tr.coefs.data(), sizeof(double) * tr.coefs.size());
        }
        return (int)MWG_OK;
    });
}

namespace {
// A tree from a complete set of node keys (every parent present) in dst's
// construction order (roots, then each split's 8 children); src_of[j] is
// the position in the key list of dst's node j.  dst is structure only.
int tree_from_keys(const MRA &mra, int n, const int *depth, const long long *l, Tree &dst, std::vector<int> &src_of) {
    Tree src;
    src.mra = &mra;
    src.structure_only = true;
    for (int i = 0; i < n; i++) {
        NodeKey k{depth[i], {l[3 * i], l[3 * i + 1], l[3 * i + 2]}};
        if (k.depth < 0 || k.depth > mra.max_depth || !mra.in_world(k.depth, k.l))
            return fail(MWG_ERR_ARG, "node key outside the world box");
        if (src.find(k.depth, k.l) >= 0) return fail(MWG_ERR_ARG, "duplicate node key");
        src.add_node(k, -1);
    }
    dst.mra = &mra;
    dst.structure_only = true;
    dst.make_roots();
    src_of.assign(dst.size(), -1);
    for (int i = 0; i < dst.size(); i++) {
        int j = src.find(dst.keys[i].depth, dst.keys[i].l);
        if (j < 0) return fail(MWG_ERR_ARG, "root node missing");
        src_of[i] = j;
    }
    for (int i = 0; i < dst.size(); i++) {
        const NodeKey k = dst.keys[i];
        long long c0[3] = {2 * k.l[0], 2 * k.l[1], 2 * k.l[2]};
        if (src.find(k.depth + 1, c0) < 0) continue;
        int first = dst.split(i);
        src_of.resize(dst.size(), -1);
        for (int c = 0; c < 8; c++) {
            int j = src.find(dst.keys[first + c].depth, dst.keys[first + c].l);
            if (j < 0) return fail(MWG_ERR_ARG, "incomplete node family (a child is missing)");
            src_of[first + c] = j;
        }
    }
    if (dst.size() != n) return fail(MWG_ERR_ARG, "nodes unreachable from the roots (missing parent)");
    dst.structure_only = false;
    return MWG_OK;
}
}  // namespace

int mwg_tree_set(mwg_tree t, int n, const int *depth, const long long *l, const double *coefs) {
    return guarded([&] {
        if (!t || n < 1 || !depth || !l) return fail(MWG_ERR_ARG, "bad arguments");
        const MRA &mra = *t->tree.mra;
        const int blk = mra.block();
        Tree dst;
        std::vector<int> src_of;
        if (int rc = tree_from_keys(mra, n, depth, l, dst, src_of)) return rc;
        dst.coefs.resize((size_t)n * blk);  // every block is written below
#pragma omp parallel for num_threads(team_size()) schedule(static) if (n > 4096)
        for (int i = 0; i < n; i++) {
            if (coefs)
                std::memcpy(dst.node(i), coefs + (size_t)src_of[i] * blk, sizeof(double) * blk);
            else
                std::fill_n(dst.node(i), blk, 0.0);
        }
        dst.sqnorm = tree_square_norm(dst);
        t->tree = std::move(dst);
        t->tree.mra = &mra;
        host_changed(t);
        t->apply_thr.clear();
        return MWG_OK;
    });
}

// Packed tree (wire and file format), little-endian:
//   header (72 bytes): magic "MWGTREE1", int32 version (1), kp1, basis type,
//     root scale, corner[3], boxes[3], double scaling factor, int64 node
//     count n, double squared norm;
//   int32 depth[n], zero padding to a multiple of 8 bytes;
//   int64 l[n][3];
//   double coefs[n][8 (k+1)^3] (compressed blocks);
// nodes in tree order (every parent before its children).
namespace {
struct PackedHeader {
    char magic[8];
    int32_t version, kp1, basis, n0;
    int32_t corner[3], boxes[3];
    double sfac;
    int64_t n_nodes;
    double sqnorm;
};
static_assert(sizeof(PackedHeader) == 72, "packed header layout");
constexpr char kPackMagic[8] = {'M', 'W', 'G', 'T', 'R', 'E', 'E', '1'};

size_t pad8(size_t b) { return (b + 7) / 8 * 8; }
size_t packed_bytes(long long n, int blk) {
    return sizeof(PackedHeader) + pad8(sizeof(int32_t) * (size_t)n) + sizeof(int64_t) * 3 * (size_t)n +
           sizeof(double) * (size_t)n * blk;
}
}  // namespace

int mwg_tree_pack_size(mwg_tree t, long long *bytes) {
    if (!t || !bytes) return fail(MWG_ERR_ARG, "null pointer argument");
    *bytes = (long long)packed_bytes(t->tree.size(), t->tree.mra->block());
    return MWG_OK;
}

namespace {
// The packed form's header and keys (everything before the blocks) into p.
void pack_head(const Tree &T, char *p) {
    const MRA &mra = *T.mra;
    const long long n = T.size();
    PackedHeader h{};
        std::memcpy(h.magic, kPackMagic, 8);
        h.version = 1;
        h.kp1 = mra.kp1();
        h.basis = mra.basis.type;
        h.n0 = mra.n0;
        for (int d = 0; d < 3; d++) {
            h.corner[d] = mra.corner[d];
            h.boxes[d] = mra.boxes[d];
        }
        h.sfac = mra.sfac;
        h.n_nodes = n;
        h.sqnorm = T.sqnorm;
        std::memcpy(p, &h, sizeof(h));
        p += sizeof(h);
        std::vector<int32_t> dep(n);
        std::vector<int64_t> l(3 * n);
        for (long long i = 0; i < n; i++) {
            dep[i] = T.keys[i].depth;
            for (int d = 0; d < 3; d++) l[3 * i + d] = T.keys[i].l[d];
        }
        const size_t db = sizeof(int32_t) * (size_t)n;
        std::memcpy(p, dep.data(), db);
        std::memset(p + db, 0, pad8(db) - db);
        p += pad8(db);
        std::memcpy(p, l.data(), sizeof(int64_t) * 3 * (size_t)n);
}
size_t head_bytes(long long n) { return sizeof(PackedHeader) + pad8(sizeof(int32_t) * (size_t)n) + sizeof(int64_t) * 3 * (size_t)n; }

// Validates a packed header against mra; the node count on success.
long long check_head(const MRA &mra, const PackedHeader &h, long long bytes) {
    if (std::memcmp(h.magic, kPackMagic, 8) != 0) throw std::invalid_argument("packed tree: bad magic");
    if (h.version != 1) throw std::invalid_argument("packed tree: unknown version");
    bool same = h.kp1 == mra.kp1() && h.basis == mra.basis.type && h.n0 == mra.n0 &&
                std::memcmp(&h.sfac, &mra.sfac, sizeof(double)) == 0;
    for (int d = 0; d < 3; d++) same = same && h.corner[d] == mra.corner[d] && h.boxes[d] == mra.boxes[d];
    if (!same) throw std::invalid_argument("packed tree: written on a different MRA");
    if (h.n_nodes < 1 || h.n_nodes >= (1LL << 31)) throw std::invalid_argument("packed tree: bad node count");
    if ((size_t)bytes != packed_bytes(h.n_nodes, mra.block())) throw std::invalid_argument("packed tree: size mismatch");
    return h.n_nodes;
}
}  // namespace

int mwg_tree_pack(mwg_tree t, void *buffer, long long bytes) {
    return guarded([&] {
        if (!t || !buffer) return fail(MWG_ERR_ARG, "null pointer argument");
        ensure_host(t);
        const Tree &T = t->tree;
        const long long n = T.size();
        if (bytes < 0 || (size_t)bytes < packed_bytes(n, T.mra->block())) return fail(MWG_ERR_ARG, "buffer too small");
        char *p = static_cast<char *>(buffer);
        pack_head(T, p);
        std::memcpy(p + head_bytes(n), T.coefs.data(), sizeof(double) * T.coefs.size());
        return (int)MWG_OK;
    });
}

int mwg_tree_pack_device(mwg_tree t, mwg_context ctx, void *dst, long long bytes) {
    return guarded_ctx(ctx, [&] {
        if (!t || !ctx || !dst) return fail(MWG_ERR_ARG, "null pointer argument");
        const Tree &T = t->tree;
        const long long n = T.size();
        if (bytes < 0 || (size_t)bytes < packed_bytes(n, T.mra->block())) return fail(MWG_ERR_ARG, "buffer too small");
        std::vector<char> head(head_bytes(n));
        pack_head(T, head.data());
        std::string err;
        // the head as doubles: its size is a multiple of 8 bytes
        dev_check(dev_upload_into(ctx->dev, reinterpret_cast<const double *>(head.data()), head.size() / 8,
                                  static_cast<double *>(dst), err),
                  err);
        TempBlocks tb{ctx->dev, {}};
        const double *blocks = tb.of(ctx, t);
        dev_check(dev_d2d(ctx->dev, reinterpret_cast<double *>(static_cast<char *>(dst) + head.size()), blocks,
                          tree_values(t), err),
                  err);
        return (int)MWG_OK;
    });
}

int mwg_device_buffer_create(mwg_context ctx, long long bytes, void **dev) {
    return guarded_ctx(ctx, [&] {
        if (!ctx || !dev || bytes < 0) return fail(MWG_ERR_ARG, "bad arguments");
        std::string err;
        double *p = nullptr;
        dev_check(dev_alloc(ctx->dev, ((size_t)bytes + 7) / 8, &p, err), err);
        *dev = p;
        return (int)MWG_OK;
    });
}

int mwg_device_buffer_destroy(mwg_context ctx, void *dev) {
    if (!ctx) return fail(MWG_ERR_ARG, "null context");
    dev_free(ctx->dev, dev);
    return MWG_OK;
}

int mwg_device_buffer_read(mwg_context ctx, const void *dev, void *host, long long bytes) {
    return guarded_ctx(ctx, [&] {
        if (!ctx || !dev || !host || bytes < 0) return fail(MWG_ERR_ARG, "bad arguments");
        std::string err;
        dev_check(dev_copy_d2h(ctx->dev, host, dev, (size_t)bytes, err), err);
        return (int)MWG_OK;
    });
}

int mwg_tree_unpack_device(mwg_tree t, mwg_context ctx, const void *src, long long bytes) {
    return guarded_ctx(ctx, [&] {
        if (!t || !ctx || !src) return fail(MWG_ERR_ARG, "null pointer argument");
        if (t->plan_refs > 0) return fail(MWG_ERR_ARG, "unpack: a plan still writes into this tree");
        const MRA &mra = *t->tree.mra;
        if (bytes < (long long)sizeof(PackedHeader)) return fail(MWG_ERR_ARG, "packed tree: truncated header");
        std::string err;
        PackedHeader h;
        dev_check(dev_download(ctx->dev, static_cast<const double *>(src), sizeof(h) / 8, reinterpret_cast<double *>(&h),
                               err),
                  err);
        const long long n = check_head(mra, h, bytes);
        std::vector<char> head(head_bytes(n));
        dev_check(dev_download(ctx->dev, static_cast<const double *>(src), head.size() / 8,
                               reinterpret_cast<double *>(head.data()), err),
                  err);
        std::vector<int> dep(n);
        std::vector<long long> l(3 * n);
        {
            const char *p = head.data() + sizeof(h);
            std::vector<int32_t> d32(n);
            std::memcpy(d32.data(), p, sizeof(int32_t) * (size_t)n);
            for (long long i = 0; i < n; i++) dep[i] = d32[i];
            p += pad8(sizeof(int32_t) * (size_t)n);
            std::vector<int64_t> l64(3 * n);
            std::memcpy(l64.data(), p, sizeof(int64_t) * 3 * (size_t)n);
            for (long long i = 0; i < 3 * n; i++) l[i] = l64[i];
        }
        Tree dst;
        std::vector<int> src_of;
        if (int rc = tree_from_keys(mra, (int)n, dep.data(), l.data(), dst, src_of)) return rc;
        // the tree becomes resident on ctx, its blocks gathered on the device
        if (t->dctx != ctx) {
            if (int rc = mwg_tree_set_resident(t, ctx, 1)) return rc;
        }
        const int blk = mra.block();
        DeviceResult res{ctx->dev, nullptr};
        dev_check(dev_alloc(ctx->dev, (size_t)n * blk, &res.p, err), err);
        const double *sb = reinterpret_cast<const double *>(static_cast<const char *>(src) + head.size());
        dev_check(dev_gather_blocks(ctx->dev, sb, src_of, blk, res.p, err), err);
        std::vector<double> s2(n), w2(n);
        dev_check(dev_node_norms(ctx->dev, mra.kp1(), res.p, n, s2.data(), w2.data(), err), err);
        t->tree = std::move(dst);
        t->tree.mra = &mra;
        t->apply_thr.clear();
        take_device_result(ctx, t, res.release(), std::move(s2), std::move(w2));
        return (int)MWG_OK;
    });
}

int mwg_tree_unpack(mwg_tree t, const void *buffer, long long bytes) {
    return guarded([&] {
        if (!t || !buffer) return fail(MWG_ERR_ARG, "null pointer argument");
        const MRA &mra = *t->tree.mra;
        if (bytes < (long long)sizeof(PackedHeader)) return fail(MWG_ERR_ARG, "packed tree: truncated header");
        PackedHeader h;
        std::memcpy(&h, buffer, sizeof(h));
        const long long n = check_head(mra, h, bytes);
        const char *p = static_cast<const char *>(buffer) + sizeof(h);
        std::vector<int> dep(n);
        std::vector<int32_t> d32(n);
        std::memcpy(d32.data(), p, sizeof(int32_t) * (size_t)n);
        for (long long i = 0; i < n; i++) dep[i] = d32[i];
        p += pad8(sizeof(int32_t) * (size_t)n);
        std::vector<long long> l(3 * n);
        std::vector<int64_t> l64(3 * n);
        std::memcpy(l64.data(), p, sizeof(int64_t) * 3 * (size_t)n);
        for (long long i = 0; i < 3 * n; i++) l[i] = l64[i];
        p += sizeof(int64_t) * 3 * (size_t)n;
        std::vector<double> c((size_t)n * mra.block());
        std::memcpy(c.data(), p, sizeof(double) * c.size());
        return mwg_tree_set(t, (int)n, dep.data(), l.data(), c.data());
    });
}

int mwg_tree_save(mwg_tree t, const char *path) {
    return guarded([&] {
        if (!t || !path) return fail(MWG_ERR_ARG, "null pointer argument");
        long long nb = 0;
        int rc = mwg_tree_pack_size(t, &nb);
        if (rc) return rc;
        std::vector<char> buf((size_t)nb);
        rc = mwg_tree_pack(t, buf.data(), nb);
        if (rc) return rc;
        FILE *f = std::fopen(path, "wb");
        if (!f) return fail(MWG_ERR_RUNTIME, std::string("cannot open ") + path + " for writing");
        const size_t w = std::fwrite(buf.data(), 1, buf.size(), f);
        const int cl = std::fclose(f);
        if (w != buf.size() || cl != 0) return fail(MWG_ERR_RUNTIME, std::string("write error on ") + path);
        return (int)MWG_OK;
    });
}

int mwg_tree_load(mwg_tree t, const char *path) {
    return guarded([&] {
        if (!t || !path) return fail(MWG_ERR_ARG, "null pointer argument");
        FILE *f = std::fopen(path, "rb");
        if (!f) return fail(MWG_ERR_RUNTIME, std::string("cannot open ") + path);
        std::vector<char> buf;
        char tmp[1 << 16];
        size_t r;
        while ((r = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + r);
        const bool err = std::ferror(f) != 0;
        std::fclose(f);
        if (err) return fail(MWG_ERR_RUNTIME, std::string("read error on ") + path);
        return mwg_tree_unpack(t, buf.data(), (long long)buf.size());
    });
}

int mwg_tree_square_norm(mwg_tree t, double *s) {
    if (!t || !s) return fail(MWG_ERR_ARG, "null pointer argument");
    // a tree whose blocks only the device holds: the norm its last device
    // call computed
    *s = t->host_valid ? tree_square_norm(t->tree) : t->tree.sqnorm;
    return MWG_OK;
}

int mwg_tree_dot(mwg_tree a, mwg_tree b, double *s) {
    return guarded([&] {
        if (!a || !b || !s) return fail(MWG_ERR_ARG, "null pointer argument");
        if (a->tree.mra != b->tree.mra) return fail(MWG_ERR_ARG, "trees on different MRAs");
        ensure_host(a);
        ensure_host(b);
        *s = tree_dot(a->tree, b->tree);
        return (int)MWG_OK;
    });
}

int mwg_tree_eval(mwg_tree t, int n, const double *r, double *v) {
    return guarded([&] {
        if (!t || n < 0 || (n > 0 && (!r || !v))) return fail(MWG_ERR_ARG, "bad arguments");
        ensure_host(t);
        for (int i = 0; i < n; i++) v[i] = tree_eval(t->tree, r + 3 * i);
        return (int)MWG_OK;
    });
}

int mwg_tree_rescale(mwg_tree t, double f) {
    return guarded([&] {
        if (!t) return fail(MWG_ERR_ARG, "null tree");
        if (!std::isfinite(f)) return fail(MWG_ERR_ARG, "non-finite factor");
        if (t->host_valid)
            for (double &x : t->tree.coefs) x *= f;
        if (t->dev_valid) {
            std::string err;
            dev_check(dev_scale(t->dctx->dev, t->dblocks, tree_values(t), f, err), err);
        }
        t->tree.sqnorm *= f * f;
        for (double &x : t->s2) x *= f * f;
        for (double &x : t->w2) x *= f * f;
        for (double &x : t->cn8) x *= std::fabs(f);
        return (int)MWG_OK;
    });
}

int mwg_tree_thresholds(mwg_tree t, double *thr) {
    if (!t || !thr) return fail(MWG_ERR_ARG, "null pointer argument");
    if (t->apply_thr.size() != (size_t)t->tree.size()) return fail(MWG_ERR_ARG, "tree holds no apply result");
    std::memcpy(thr, t->apply_thr.data(), sizeof(double) * t->apply_thr.size());
    return MWG_OK;
}

int mwg_apply(mwg_context ctx, double prec, mwg_tree out, mwg_operator op, mwg_tree inp, int max_iter, int abs_prec) {
    return mwg_apply_prec_trees(ctx, prec, out, op, inp, nullptr, 0, max_iter, abs_prec);
}

int mwg_apply_prec_trees(mwg_context ctx, double prec, mwg_tree out, mwg_operator op, mwg_tree inp,
                         const mwg_tree *prec_trees, int n_prec, int max_iter, int abs_prec) {
    mwg_apply_desc d;
    d.out = out;
    d.op = op;
    d.inp = inp;
    d.prec = prec;
    d.max_iter = max_iter;
    d.abs_prec = abs_prec;
    d.prec_trees = prec_trees;
    d.n_prec = n_prec;
    return mwg_apply_batch(ctx, 1, &d);
}

int mwg_apply_batch(mwg_context ctx, int n, const mwg_apply_desc *desc) {
    return guarded_ctx(ctx, [&] {
        if (!ctx || n < 0 || (n > 0 && !desc)) return fail(MWG_ERR_ARG, "bad arguments");
        std::vector<Job> jobs(n);
        for (int i = 0; i < n; i++) {
            const mwg_apply_desc &d = desc[i];
            if (!d.out || !d.op || !d.inp) return fail(MWG_ERR_ARG, "null handle");
            if (d.out == d.inp) return fail(MWG_ERR_ARG, "apply: output and input must be different trees");
            if (d.n_prec < 0 || (d.n_prec > 0 && !d.prec_trees)) return fail(MWG_ERR_ARG, "bad prec tree list");
            Job &j = jobs[i];
            j.out = d.out;
            j.op = d.op;
            j.inp = &d.inp->tree;
            j.inp_h = d.inp;
            j.prec = d.prec;
            j.max_iter = d.max_iter;
            j.abs_prec = d.abs_prec != 0;
            for (int t = 0; t < d.n_prec; t++) {
                if (!d.prec_trees[t]) return fail(MWG_ERR_ARG, "null prec tree");
                j.prec_h.push_back(d.prec_trees[t]);
            }
        }
        return apply_batch(ctx, jobs);
    });
}

int mwg_apply_fixed(mwg_context ctx, mwg_tree out, mwg_operator op, mwg_tree inp, const double *thr, int mode) {
    return guarded_ctx(ctx, [&] {
        mwg_plan plan = nullptr;
        int rc = make_plan(ctx, 1, &out, &op, &inp, &thr, mode, &plan);
        if (rc) return rc;
        std::unique_ptr<mwg_plan_s, int (*)(mwg_plan)> hold(plan, mwg_plan_destroy);
        std::string err;
        double ms = 0.0;
        unsigned long long apps = 0;
        rc = dev_plan_run(plan->dev, mode, plan->tau, &ms, &apps, nullptr, err, nullptr);
        if (rc) return dev_fail(rc, err);
        rc = plan_finish(plan);
        if (rc) return rc;
        auto &st = ctx->last_stats;
        st.assign(8, 0.0);
        st[0] = ms;
        st[1] = plan->total_out;
        st[2] = plan->slots;
        st[3] = plan->gen;
        st[4] = plan->plan_ms;
        st[5] = 1;
        st[6] = (double)apps;
        return MWG_OK;
    });
}

int mwg_plan_create(mwg_context ctx, mwg_tree out, mwg_operator op, mwg_tree inp, const double *thr, int mode,
                    mwg_plan *plan) {
    return guarded_ctx(ctx, [&] {
        if (out == inp) return fail(MWG_ERR_ARG, "apply: output and input must be different trees");
        return make_plan(ctx, 1, &out, &op, &inp, &thr, mode, plan);
    });
}

int mwg_plan_create_batch(mwg_context ctx, int n, const mwg_tree *out, const mwg_operator *op, const mwg_tree *inp,
                          const double *const *thr, int mode, mwg_plan *plan) {
    return guarded_ctx(ctx, [&] { return make_plan(ctx, n, out, op, inp, thr, mode, plan); });
}

int mwg_plan_execute(mwg_plan plan, double *stats, int n) {
    return guarded_ctx(plan ? plan->ctx : nullptr, [&] {
        if (!plan) return fail(MWG_ERR_ARG, "null plan");
        std::string err;
        double ms = 0.0, split[2] = {0.0, 0.0};
        unsigned long long apps = 0, lp = 0;
        plan->post_done = false;
        int rc = dev_plan_run(plan->dev, plan->mode, plan->tau, &ms, &apps, &lp, err, split, false, nullptr,
                              plan->use_graph);
        if (rc) return dev_fail(rc, err);
        const double v[5] = {ms, (double)apps, (double)lp, split[0], split[1]};
        for (int i = 0; stats && i < n && i < 5; i++) stats[i] = v[i];
        return MWG_OK;
    });
}

int mwg_plan_apply(mwg_plan plan, double *stats, int n) {
    return guarded_ctx(plan ? plan->ctx : nullptr, [&] {
        if (!plan) return fail(MWG_ERR_ARG, "null plan");
        std::string err;
        double ms = 0.0, split[2] = {0.0, 0.0}, pms = 0.0;
        unsigned long long apps = 0, lp = 0;
        plan->post_done = false;
        int rc = dev_plan_run(plan->dev, plan->mode, plan->tau, &ms, &apps, &lp, err, split, true, &pms,
                              plan->use_graph);
        if (rc) return dev_fail(rc, err);
        plan->post_done = true;
        const double v[6] = {ms, (double)apps, (double)lp, split[0], split[1], pms};
        for (int i = 0; stats && i < n && i < 6; i++) stats[i] = v[i];
        return MWG_OK;
    });
}

int mwg_plan_tune(mwg_plan plan, double max_rel_err, int reps, double *report, int n) {
    return guarded_ctx(plan ? plan->ctx : nullptr, [&] {
        if (!plan) return fail(MWG_ERR_ARG, "null plan");
        if (!(max_rel_err > 0.0) || !std::isfinite(max_rel_err) || reps < 1)
            return fail(MWG_ERR_ARG, "tune: max_rel_err must be > 0 and reps >= 1");
        std::string err;
        DevicePlan *dp = plan->dev;
        const bool lowp = plan->mode != MWG_MODE_FP64;
        // FP64 reference of the plan on the device
        double ms = 0.0;
        unsigned long long apps = 0;
        int rc = dev_plan_run(dp, MWG_MODE_FP64, plan->tau, &ms, &apps, nullptr, err);
        double *ref = nullptr;
        if (!rc) rc = dev_plan_snapshot(dp, &ref, err);
        if (rc) return dev_fail(rc, err);
        struct Free {
            DeviceCtx *c;
            double *p;
            ~Free() { dev_free(c, p); }
        } hold{plan->ctx->dev, ref};
        // one configuration: its error against the reference (largest
        // difference over the largest value) and its mean device time
        int tried = 0;
        auto measure = [&](const TunedKnobs &k, double &t_ms, double &rel) -> int {
            int r = dev_plan_set_knobs(dp, k.single_ratio, k.order, err);
            if (!r) r = dev_plan_run(dp, plan->mode, k.tau, &ms, &apps, nullptr, err);
            double md = 0.0, mr = 0.0;
            if (!r) r = dev_plan_max_diff(dp, ref, &md, &mr, err);
            if (r) return r;
            rel = mr > 0.0 ? md / mr : 0.0;
            t_ms = 0.0;
            for (int q = 0; q < reps && !r; q++) {
                r = dev_plan_run(dp, plan->mode, k.tau, &ms, &apps, nullptr, err);
                t_ms += ms / reps;
            }
            tried++;
            return r;
        };
        TunedKnobs start;
        start.tau = plan->tau;
        dev_plan_knobs(dp, &start.single_ratio, &start.order);
        TunedKnobs best = start;
        double best_ms = 0.0, best_err = 0.0;
        rc = measure(best, best_ms, best_err);
        if (rc) return dev_fail(rc, err);
        const double base_ms = best_ms, base_err = best_err;
        // a starting setting outside the bound is replaced by any that meets it
        bool found = best_err <= max_rel_err;
        if (!found) best_ms = HUGE_VAL;
        // coordinate search: tau (low-precision modes), then the single-
        // rounding cut (split mode), then the queue order; a configuration
        // counts only if its error stays within max_rel_err
        auto consider = [&](const TunedKnobs &k) -> int {
            double t = 0.0, e = 0.0;
            int r = measure(k, t, e);
            if (r) return r;
            if (e <= max_rel_err && t < best_ms) {
                best = k;
                best_ms = t;
                best_err = e;
                found = true;
            }
            return 0;
        };
        if (lowp) {
            for (double tau : {4096.0, 16384.0, 65536.0, 262144.0, 1048576.0, 4194304.0}) {
                if (tau == best.tau) continue;
                TunedKnobs k = best;
                k.tau = tau;
                if ((rc = consider(k))) return dev_fail(rc, err);
            }
        }
        if (plan->mode == MWG_MODE_SPLIT16) {
            for (double sr : {0.0, 4.0, 16.0, 64.0, 256.0}) {
                if (sr == best.single_ratio) continue;
                TunedKnobs k = best;
                k.single_ratio = sr;
                if ((rc = consider(k))) return dev_fail(rc, err);
            }
        }
        {
            TunedKnobs k = best;
            k.order = 1 - best.order;
            if ((rc = consider(k))) return dev_fail(rc, err);
        }
        if (!found) {
            rc = dev_plan_set_knobs(dp, start.single_ratio, start.order, err);
            if (rc) return dev_fail(rc, err);
            return fail(MWG_ERR_ARG, "tune: no setting keeps the error within max_rel_err");
        }
        rc = dev_plan_set_knobs(dp, best.single_ratio, best.order, err);
        if (rc) return dev_fail(rc, err);
        plan->tau = best.tau;
        plan->post_done = false;
        plan->ctx->tuned[plan->tune_key] = best;
        const double v[9] = {best.tau, best.single_ratio, (double)best.order, best_ms, base_ms, best_err,
                             (double)tried, best_ms > 0.0 ? base_ms / best_ms : 0.0, base_err};
        for (int i = 0; report && i < n && i < 9; i++) report[i] = v[i];
        return MWG_OK;
    });
}

int mwg_plan_use_graph(mwg_plan plan, int on) {
    if (!plan) return fail(MWG_ERR_ARG, "null plan");
    plan->use_graph = on != 0;
    return MWG_OK;
}

int mwg_plan_finish(mwg_plan plan) {
    return guarded_ctx(plan ? plan->ctx : nullptr, [&] {
        if (!plan) return fail(MWG_ERR_ARG, "null plan");
        return plan_finish(plan);
    });
}

int mwg_plan_download_ns(mwg_plan plan, double *ns, long long capacity) {
    return guarded_ctx(plan ? plan->ctx : nullptr, [&] {
        if (!plan || !ns) return fail(MWG_ERR_ARG, "null pointer argument");
        if (plan->post_done)
            return fail(MWG_ERR_ARG, "download_ns: the last execution ran the post transforms (mwg_plan_apply); "
                                     "the raw non-standard blocks come from mwg_plan_execute");
        const long long need = (long long)plan->total_out * plan->outs[0]->tree.mra->block();
        if (capacity < need)
            return fail(MWG_ERR_ARG, "download_ns: the buffer holds " + std::to_string(capacity) + " values, the plan has " +
                                         std::to_string(need));
        std::string err;
        int rc = dev_plan_download(plan->dev, ns, err);
        if (rc) return dev_fail(rc, err);
        return MWG_OK;
    });
}

int mwg_plan_info(mwg_plan plan, double *info, int n) {
    if (!plan || !info) return fail(MWG_ERR_ARG, "null pointer argument");
    double sr = 0.0;
    int order = 0;
    dev_plan_knobs(plan->dev, &sr, &order);
    double v[11] = {(double)plan->total_out, plan->slots, plan->gen, plan->plan_ms, plan->host_apps, plan->segs,
                    (double)plan->outs.size(), plan->band_w, plan->tau, sr, (double)order};
    for (int i = 0; i < n && i < 11; i++) info[i] = v[i];
    return MWG_OK;
}

int mwg_plan_destroy(mwg_plan plan) {
    if (!plan) return MWG_OK;
    dev_plan_destroy(plan->dev);
    delete plan;
    return MWG_OK;
}

int mwg_context_set_precision(mwg_context ctx, int mode, double tau) {
    if (!ctx) return fail(MWG_ERR_ARG, "null context");
    if (mode != MWG_MODE_FP64 && mode != MWG_MODE_MIXED && mode != MWG_MODE_SPLIT16)
        return on_ctx(ctx, fail(MWG_ERR_ARG, "unknown arithmetic mode"));
    if (!(tau >= 1.0) || !std::isfinite(tau))
        return on_ctx(ctx, fail(MWG_ERR_ARG, "mixed-precision ratio must be finite and >= 1"));
    ctx->mode = mode;
    ctx->tau = tau;
    return MWG_OK;
}

int mwg_last_apply_stats(mwg_context ctx, double *stats, int n) {
    if (!ctx || !stats) return fail(MWG_ERR_ARG, "null pointer argument");
    for (int i = 0; i < n && i < (int)ctx->last_stats.size(); i++) stats[i] = ctx->last_stats[i];
    return MWG_OK;
}

}  // extern "C"
