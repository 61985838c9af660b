// Interface between the host driver and the HIP apply kernels.
#pragma once

#include <string>
#include <vector>

#include "mw_internal.hpp"

namespace mwg {

struct DeviceCtx;

// Returns 0 on success; on failure a negative code and a message in err
// (-1: invalid argument, -2: device error).
int dev_create(int device, DeviceCtx **out, std::string &err);
void dev_destroy(DeviceCtx *c);
// Drop the device copy of an operator's tables (called when it is destroyed).
void dev_forget_operator(DeviceCtx *c, const void *key);

// One apply of a (possibly batched) device plan: the operator (its host
// tables, keyed for the context's resident copy) and the host plan.
struct PlanPart {
    const void *op_key;
    const OperFlat *flat;
    const ApplyPlan *plan;
    // the plan's pool already on the device (dev_pool_fill), copied into the
    // device plan; otherwise plan->in_coefs is uploaded
    const double *dev_pool = nullptr;
};

// A plan resident in device memory: band tables, input slots and output
// blocks of one or more applies, run together in each launch.
struct DevicePlan;
// adopt_pool: the parts' device pools lie one after another in this
// allocation, in part order; the plan takes it as its pool (no copy) and
// frees it.  Without it each part's pool is copied (or uploaded).
int dev_plan_create(DeviceCtx *c, const std::vector<PlanPart> &parts, DevicePlan **out, std::string &err,
                    double *adopt_pool = nullptr);
// Run the per-node calculator over the resident plan.  kernel_ms receives the
// device time of the launches, term_apps the number of term-applications that
// passed screening.
// mode 0: every term in FP64.  mode 1: terms with bound > tau * thr in FP64,
// the rest on FP32 MFMA, their sum added to the FP64 output.  lp_apps
// receives the number of FP32 term-applications.
// split_ms (optional, 2 values): device time of the FP64-class launch and of
// the low-precision launch (they run concurrently, on two streams).
// With post (after dev_plan_set_post) the post transforms and the node
// norms run in place on the output blocks after the calculator, inside the
// timed region; post_ms receives their device time.
// With graph the step (zeroing, class launches, post) replays a hipGraph
// captured on the plan for (mode, tau, post) after one eager run; a replay
// reports the whole device time only (split_ms and post_ms stay 0).
int dev_plan_run(DevicePlan *p, int mode, double tau, double *kernel_ms, unsigned long long *term_apps,
                 unsigned long long *lp_apps, std::string &err, double *split_ms = nullptr, bool post = false,
                 double *post_ms = nullptr, bool graph = false);
// Post transforms of the plan's output grid: levels[d] holds the (branch
// node, first child) pairs at depth d as indices into the plan's output
// nodes.  Uploads the pairs with their device addresses and the filter.
int dev_plan_set_post(DevicePlan *p, const MRA &mra, const std::vector<std::vector<std::pair<long long, long long>>> &levels,
                      std::string &err);
// Tunable knobs of a plan: the split class's single-rounding cut (items with
// bound <= single_ratio x thr_cls round their intermediates once) and the
// work-queue order (0: segment order, 1: largest passing-term count first).
int dev_plan_set_knobs(DevicePlan *p, double single_ratio, int order, std::string &err);
void dev_plan_knobs(const DevicePlan *p, double *single_ratio, int *order);
// A device copy of the plan's output blocks (freed with dev_free), and the
// largest |output - ref| and |ref| over all output values.
int dev_plan_snapshot(DevicePlan *p, double **copy, std::string &err);
int dev_plan_max_diff(DevicePlan *p, const double *ref, double *max_diff, double *max_ref, std::string &err);
// Squared norms (scaling, wavelet) of every output node after a run with post.
int dev_plan_download_norms(DevicePlan *p, double *s2, double *w2, std::string &err);
// Copy the raw non-standard blocks of every part, in part order
// (sum of n_out x 8(k+1)^3), to the host.
int dev_plan_download(DevicePlan *p, double *out_ns, std::string &err);
// The plan's output blocks (all applies' output nodes, 8(k+1)^3 each),
// detached from the plan: the caller frees them with dev_free.
double *dev_plan_take_output(DevicePlan *p);
// A device copy of the plan's output blocks (freed with dev_free).
int dev_plan_clone_output(DevicePlan *p, double **out, std::string &err);
// Output blocks [first, first + count) of the plan's concatenated output nodes.
int dev_plan_download_range(DevicePlan *p, long long first, long long count, double *dst, std::string &err);
void dev_plan_destroy(DevicePlan *p);

// Device construction of operator tables (mw_opbuild.hip).  Each job holds an
// operator prepared by prepare_operator (terms, no tables); the build fills
// its L and norms on the host, its flat table header (f.blocks stays empty,
// f.device_blocks counts the device values) and registers the device tables
// in the context under `key`.
struct OpBuildJob {
    Operator *op;
    OperFlat *flat;
    const void *key;
};
int dev_build_operators(DeviceCtx *c, const MRA &mra, const std::vector<OpBuildJob> &jobs, std::string &err);
// Copy the device table of a device-built operator (f.device_blocks values).
int dev_download_operator_blocks(DeviceCtx *c, const void *key, double *out, long long n, std::string &err);

// Node-wise products for the multiplication (mw_multiply.hip).  For one
// output node: each input's source node (its deepest node on the path, block
// offset in the input's coefficient pool), the levels from it down to the
// output node and the child index of each level (3 bits per level, the first
// level below the source in the lowest bits), and h^{-3/2} of the output
// node's children.
struct MulNode {
    long long src_a, src_b;
    int dep_a, dep_b;
    unsigned path_a, path_b;
    double hfac;
};
struct DevMulSession;
// noalloc: every device buffer of the session is carved from the device's
// scratch array (dev_noalloc_init); no device memory is allocated, and
// inputs that do not fit are an argument error.
// a_on_device / b_on_device: that pool is device memory, used in place (or,
// with noalloc, copied into the scratch array); otherwise it is host memory
// and uploaded.
int dev_mul_begin(DeviceCtx *c, const MRA &mra, const double *pool_a, size_t n_a, const double *pool_b, size_t n_b,
                  DevMulSession **out, std::string &err, bool noalloc = false, bool a_on_device = false,
                  bool b_on_device = false);
// out_blocks: nodes.size() x 8(k+1)^3, on the host, or on the device with
// out_on_device.
int dev_mul_nodes(DevMulSession *m, const std::vector<MulNode> &nodes, double coef, double *out_blocks,
                  std::string &err, bool out_on_device = false);
void dev_mul_end(DevMulSession *m);

// The pools of plans built with device_pool (mw_multiply.hip), of every
// job in one pass (one temporary allocation, one upload, one wait):
// generated nodes reconstructed from their parents and every slot's block
// copied from the input tree's blocks on the device (inp: the input tree's
// coefficients, 8 (k+1)^3 per node in tree order) into pool
// (pl->pool_values doubles, the caller's); pl->in_norms receives the slots'
// component norms.
struct PoolFill {
    const double *inp;
    ApplyPlan *pl;
    double *pool;
};
int dev_pool_fill(DeviceCtx *c, const MRA &mra, const std::vector<PoolFill> &jobs, std::string &err);
// A device copy of n doubles (freed with dev_free).
int dev_upload(DeviceCtx *c, const double *src, size_t n, double **dst, std::string &err);
// n doubles of device memory (freed with dev_free).
int dev_alloc(DeviceCtx *c, size_t n, double **dst, std::string &err);
// Block i (blk doubles) of src to block dst_idx[i] of dst, on the device.
int dev_scatter_blocks(DeviceCtx *c, const double *src, const std::vector<int> &dst_idx, int blk, double *dst,
                       std::string &err);
// Block src_idx[i] of src to block i of dst, on the device.
int dev_gather_blocks(DeviceCtx *c, const double *src, const std::vector<int> &src_idx, int blk, double *dst,
                      std::string &err);
// p[0, n) *= f on the device, finished on return.
int dev_scale(DeviceCtx *c, double *p, size_t n, double f, std::string &err);
// n doubles from host memory into device memory at dst, finished on return.
int dev_upload_into(DeviceCtx *c, const double *src, size_t n, double *dst, std::string &err);
// n doubles from device memory to device memory, finished on return.
int dev_d2d(DeviceCtx *c, double *dst, const double *src, size_t n, std::string &err);

// Node blocks resident on the device (mw_multiply.hip).  Per node of
// blocks[0, n): its scaling part's and wavelet parts' squared norms.
int dev_node_norms(DeviceCtx *c, int kp1, const double *blocks, long long n, double *host_s2, double *host_w2,
                   std::string &err);
// Per node of blocks[0, n): the norms of its 8 components (8 per node).
int dev_component_norms(DeviceCtx *c, int kp1, const double *blocks, long long n, double *host8, std::string &err);
// A branch node's block and its 8 children's blocks (consecutive).
struct PostPair {
    double *node, *child0;
};
// Post transforms of an NS output tree whose blocks are on the device, the
// branch nodes grouped by depth (levels[d]): TopDown (each branch's
// reconstruction added into its children's scaling parts, shallow to deep),
// then BottomUp (each branch's block recomputed from its children's scaling
// parts, deep to shallow).
// Without topdown only the BottomUp steps run (a compressed tree from its
// leaves' scaling parts).
int dev_post_transform(DeviceCtx *c, const MRA &mra, const std::vector<std::vector<PostPair>> &levels,
                       std::string &err, bool topdown = true);
// The same TopDown/BottomUp sequence over pairs already on the device:
// level d's pairs are pairs[lvl_off[d], lvl_off[d + 1]); F is the two-scale
// filter on the device.  Launches only, on `stream`.
int dev_post_launch(void *stream, int kp1, const PostPair *pairs, const std::vector<long long> &lvl_off,
                    const double *F, std::string &err, bool topdown = true);
// Squared norms of the scaling and wavelet parts of blocks[0, n) into the
// device arrays s2, w2 (a launch on `stream`).
int dev_node_norms_launch(void *stream, int kp1, const double *blocks, long long n, double *s2, double *w2,
                          std::string &err);
int dev_download(DeviceCtx *c, const double *src, size_t n, double *dst, std::string &err);
// Device-to-host copy on the context's stream, finished on return; large
// copies go through pinned staging buffers (mw_device.hip).
int dev_copy_d2h(DeviceCtx *c, void *dst, const void *src, size_t bytes, std::string &err);
// Host-to-device copy on the context's stream, finished on return; large
// copies go through the same pinned staging buffers.
int dev_copy_h2d(DeviceCtx *c, void *dst, const void *src, size_t bytes, std::string &err);
void dev_free(DeviceCtx *c, void *p);

// Node-wise sums (mw_multiply.hip): for one output node and one input, the
// input's source node (its deepest node on the path: block offset in the
// concatenated pool of all inputs, or -1 when the input has no node on the
// path), the levels from it down to the output node and the child index of
// each level (3 bits per level, the first level below the source lowest).
struct AddSrc {
    long long off;
    int dep, pad_;
    unsigned long long path;
};
constexpr int kAddMaxLevels = 21;  // levels a 64-bit path holds
struct DevAddSession;
// One input's blocks: n doubles at p, in host or device memory.
struct BlockSource {
    const double *p;
    size_t n;
    bool on_device;
};
// The inputs' blocks are concatenated in order into the session's pool.
int dev_add_begin(DeviceCtx *c, const MRA &mra, const std::vector<BlockSource> &pool, const std::vector<double> &coefs,
                  DevAddSession **out, std::string &err);
// src: n_nodes x n_inputs entries (node-major); out_blocks: n_nodes x
// 8(k+1)^3, on the host, or on the device with out_on_device.
int dev_add_nodes(DevAddSession *m, const std::vector<AddSrc> &src, double *out_blocks, std::string &err,
                  bool out_on_device = false);
void dev_add_end(DevAddSession *m);

// The device's scratch array of the allocation-free entry points
// (mw_multiply.hip): its device address, taken once per context with the
// context's device current, and stored in the context.
int dev_noalloc_init(DeviceCtx *c, std::string &err);
void dev_set_noalloc(DeviceCtx *c, void *base, size_t bytes);
void *dev_noalloc_base(DeviceCtx *c);
size_t dev_noalloc_bytes(DeviceCtx *c);

// Helpers shared by the HIP sources.
struct DevOperTables {
    int *L = nullptr;
    long long *boff = nullptr, *noff = nullptr;
    double *blocks = nullptr, *norms = nullptr;
};
int dev_device(DeviceCtx *c);
void *dev_stream(DeviceCtx *c);
// Allocate the device tables of f (blocks and norms cleared, L/boff/noff uploaded).
int dev_alloc_oper_tables(DeviceCtx *c, const OperFlat &f, DevOperTables &t, std::string &err);
// Make t the context's resident copy of operator `key` (ownership passes to
// the context).
int dev_register_operator(DeviceCtx *c, const void *key, const OperFlat &f, const DevOperTables &t, std::string &err);

}  // namespace mwg
