/* mwgreen: MI355X-native multiwavelet Green's-function apply, C ABI.
 *
 * Drop-in for the multiwavelet convolution that MRChem calls through
 *   mrcpp::apply(prec, out, ConvolutionOperator&, inp, maxIter, absPrec)
 *     src/qmoperators/two_electron/CoulombPotential.cpp:131,156
 *     src/scf_solver/HelmholtzVector.cpp:149,154
 *   mrcpp::apply(prec, out, oper, inp, precTrees, maxIter, absPrec)
 *     src/qmoperators/two_electron/ExchangePotential.cpp:200,204
 * together with the operator constructors
 *   PoissonOperator(MRA, build_prec)        src/driver.cpp:1131,1283
 *   HelmholtzOperator(MRA, mu, build_prec)  src/scf_solver/HelmholtzVector.cpp:144
 * and the FunctionTree<3> node data MRChem reads (src/mrdft/Functional.cpp:212-232).
 *
 * Conventions
 *   - All functions return MWG_OK (0) or a negative error code; the message of
 *     the last failure on the calling thread is returned by mwg_last_error().
 *     No call aborts the process (upstream MRCPP aborts through MSG_ABORT).
 *   - Node keys are (depth, lx, ly, lz): depth 0 is the root scale, and
 *     translations are absolute (root boxes have corner..corner+boxes-1).
 *   - Node blocks hold 8 (k+1)^3 doubles in compressed form: component t
 *     (bit d set = wavelet along direction d), then z, y, x with x fastest.
 *   - Handles are opaque pointers; none may be shared between threads while a
 *     call on it runs.  An apply takes `inp` and the operator read-only.
 *   - Lifetimes: a plan (mwg_plan_create*) holds device pointers into its
 *     operators' resident tables and writes into its output trees, so its
 *     context, operators and output trees must outlive it.  Destroying any of
 *     them while a plan that uses it is alive returns MWG_ERR_ARG and leaves
 *     the handle intact.  Input trees are copied when the plan is made.
 */
#ifndef MWGREEN_H
#define MWGREEN_H

#ifdef __cplusplus
extern "C" {
#endif

#define MWG_OK 0
#define MWG_ERR_ARG (-1)
#define MWG_ERR_DEVICE (-2)
#define MWG_ERR_RUNTIME (-3)

#define MWG_BASIS_INTERPOLATING 0
#define MWG_BASIS_LEGENDRE 1

/* Arithmetic of the per-node contractions.  MIXED applies every
 * term-application whose screening bound exceeds tau * threshold in FP64 and
 * the others on FP32 MFMA; SPLIT16 applies the others with FP16 hi/lo split
 * products on FP16 MFMA (hi*hi + lo*hi + hi*lo, FP32 accumulation).  Both add
 * their sums to the FP64 output; the set of terms applied is the same in all
 * modes. */
#define MWG_MODE_FP64 0
#define MWG_MODE_MIXED 1
#define MWG_MODE_SPLIT16 2

typedef struct mwg_context_s *mwg_context;
typedef struct mwg_mra_s *mwg_mra;
typedef struct mwg_operator_s *mwg_operator;
typedef struct mwg_tree_s *mwg_tree;
typedef struct mwg_plan_s *mwg_plan;

const char *mwg_last_error(void);
/* Message of the last failed call on ctx or on a plan of ctx (any thread);
 * "" before the first failure.  Valid until the next failure on ctx. */
const char *mwg_context_last_error(mwg_context ctx);

/* Device context: one HIP device, one stream, cached operator tables.
 * Replaces the OpenMP/MPI environment mrcpp::apply runs in (src/mrenv.cpp:129-135). */
int mwg_context_create(int device, mwg_context *ctx);
int mwg_context_destroy(mwg_context ctx);
/* Arithmetic used by mwg_apply / mwg_apply_prec_trees on this context
 * (default FP64; tau default 65536).  No counterpart in MRChem's calls. */
int mwg_context_set_precision(mwg_context ctx, int mode, double tau);

/* MultiResolutionAnalysis<3>: BoundingBox(min_scale, corner, boxes) with a
 * scaling factor, basis of the given order and type, and max depth.
 * Mirrors mrenv::init_mra (src/mrenv.cpp:99-127). */
int mwg_mra_create(int order, int basis_type, int root_scale, const int *corner, const int *boxes,
                   double scaling_factor, int max_depth, mwg_mra *mra);
int mwg_mra_destroy(mwg_mra mra);
/* Two-scale filter [H0 H1; G0 G1], 2(k+1) x 2(k+1) row-major. */
int mwg_mra_filter(mwg_mra mra, double *F);

/* PoissonOperator(MRA, prec): kernel 1/r (src/driver.cpp:1131). */
int mwg_poisson_create(mwg_mra mra, double build_prec, mwg_operator *op);
/* HelmholtzOperator(MRA, mu, prec): kernel exp(-mu r)/r
 * (src/scf_solver/HelmholtzVector.cpp:144). */
int mwg_helmholtz_create(mwg_mra mra, double mu, double build_prec, mwg_operator *op);
/* Operator from explicit Gaussian terms sum_m beta_m exp(-alpha_m r^2) (no
 * counterpart in MRChem's calls; used for verification). */
int mwg_gaussian_operator_create(mwg_mra mra, int n_terms, const double *alpha, const double *beta,
                                 int max_depth, mwg_operator *op);
/* Build n operators at once on the device of ctx: mu[i] = 0 gives the
 * PoissonOperator(mra, build_prec) of src/driver.cpp:1131,1283, mu[i] > 0 the
 * HelmholtzOperator(mra, mu[i], build_prec) that
 * src/scf_solver/HelmholtzVector.cpp:144 builds for every orbital in every SCF
 * iteration.  Same kernel fit and tables as mwg_poisson_create /
 * mwg_helmholtz_create (values to a few ulps); the tables stay on ctx's
 * device (mwg_operator_tables copies them out) and the operators can be used
 * only with ctx. */
int mwg_operators_create_device(mwg_context ctx, mwg_mra mra, int n, const double *mu, double build_prec,
                                mwg_operator *ops);
int mwg_operator_destroy(mwg_operator op);
int mwg_operator_info(mwg_operator op, int *n_terms, int *n_depth, int *kp1, double *r_min, double *r_max,
                      double *fit_err);
int mwg_operator_terms(mwg_operator op, double *alpha, double *beta);
/* Flattened tables: per (term m, depth d) at index m*n_depth+d the half-width
 * L (-1: empty), the offset of block (c=0, l=-L) in `blocks` and of its norm
 * in `norms`; blocks are [c][l+L][i][j] with c = 2*out_wavelet + in_wavelet. */
int mwg_operator_table_sizes(mwg_operator op, long long *n_block_values, long long *n_norm_values);
int mwg_operator_tables(mwg_operator op, int *L, long long *boff, long long *noff, double *blocks,
                        double *norms);

/* FunctionTree<3>. */
int mwg_tree_create(mwg_mra mra, mwg_tree *tree);
/* Refused (MWG_ERR_ARG) while a plan writes into the tree; frees its device
 * copy. */
int mwg_tree_destroy(mwg_tree tree);
/* Device-resident trees (MRCPP trees live in host memory; this keeps a
 * tree's blocks in HBM across calls, as the exchange chain of
 * src/qmoperators/two_electron/ExchangePotential.cpp:161-234 and the sums
 * and crops of ExchangePotentialD1.cpp:317-323,365-371 use them).
 * on = 1: the tree's blocks move to ctx's GPU at the first call on ctx that
 * reads them and stay there.  Calls on ctx read them in place (mwg_apply*,
 * mwg_multiply, mwg_add,
 * mwg_tree_add, plans) and, for a resident output, leave the result on the
 * device (built, transformed and normed there; only node norms and the grid
 * come to the host).  mwg_tree_crop and mwg_tree_rescale of a tree that only
 * the device holds run there.  Host reads (mwg_tree_get with coefficients,
 * pack, save, dot, eval) download the blocks once, lazily; host writes
 * (project, set, unpack, load) make the device copy stale until the next
 * device call uploads it.  on = 0: the host copy is made current and the
 * device copy freed.  Refused while a plan writes into the tree; a context
 * with resident trees refuses to be destroyed. */
int mwg_tree_set_resident(mwg_tree tree, mwg_context ctx, int on);
/* 1 when the tree is resident on some context, else 0; *on_host / *on_device
 * (may be NULL) say which copies are current. */
int mwg_tree_is_resident(mwg_tree tree, int *on_host, int *on_device);
/* Bytes of tree blocks moved host-to-device and device-to-host by calls on
 * ctx (inputs uploaded, results and lazy host copies downloaded) since the
 * context was made or last reset (reset = 1 zeroes them after reading). */
int mwg_context_transfer_bytes(mwg_context ctx, long long *h2d, long long *d2h, int reset);
/* Adaptive projection of sum_i c_i (a_i/pi)^{3/2} exp(-a_i |r - R_i|^2). */
int mwg_tree_project_gaussians(mwg_tree tree, double prec, int abs_prec, int max_depth, int n, const double *coef,
                               const double *exponent, const double *center);
int mwg_tree_size(mwg_tree tree, int *n_nodes);
/* Nodes in tree order (parents before children); any array may be NULL. */
int mwg_tree_get(mwg_tree tree, int *depth, long long *l, double *coefs);
/* Replace the tree by the given complete set of nodes (every parent present). */
int mwg_tree_set(mwg_tree tree, int n_nodes, const int *depth, const long long *l, const double *coefs);
int mwg_tree_square_norm(mwg_tree tree, double *sq_norm);
int mwg_tree_dot(mwg_tree a, mwg_tree b, double *dot);
int mwg_tree_eval(mwg_tree tree, int n_points, const double *r, double *values);
int mwg_tree_rescale(mwg_tree tree, double factor);
/* out = coef * a * b, built adaptively to prec (absolute with abs_prec,
 * relative to the product's norm otherwise): mrcpp::multiply(prec, out, coef,
 * a, b, max_iter, abs_prec), which MRChem calls through
 * mrcpp::cplxfunc::multiply for the pair densities rho_ij = phi_i phi_j
 * (src/qmoperators/two_electron/ExchangePotential.cpp:175) and the products
 * phi_k V_ij (ExchangePotential.cpp:213,223).  The output grid starts as the
 * union of the inputs' grids; each leaf's product is formed on the device
 * from the inputs' values at its children's points, and leaves whose wavelet
 * norm exceeds the split threshold are refined (max_iter < 0: no limit).
 * Interpolating basis only; out must differ from a and b. */
int mwg_multiply(mwg_context ctx, double prec, mwg_tree out, double coef, mwg_tree a, mwg_tree b, int max_iter,
                 int abs_prec);
/* mwg_multiply with options.  MWG_MUL_MAX_NORMS: the useMaxNorms refinement
 * that every cplxfunc::multiply call of the exchange passes
 * (ExchangePotential.cpp:175,213,223, last argument true).  The grid grows
 * from the roots; a node splits when maxW_a maxS_b + maxW_b maxS_a > prec,
 * with maxS and maxW the square roots of the largest 2^{3n} ||block||^2 and
 * 2^{3n} ||wavelet part||^2 over the input's node and its descendants (n the
 * node's scale; below an input's finest node its scaling value with maxW = 0).
 * abs_prec is unused with it (the rule is absolute). */
#define MWG_MUL_MAX_NORMS 1
int mwg_multiply_ex(mwg_context ctx, double prec, mwg_tree out, double coef, mwg_tree a, mwg_tree b, int max_iter,
                    int abs_prec, int flags);
/* mwg_multiply that allocates no device memory: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight per-pair loops, harnesses that forbid hipMalloc).  Every device
 * buffer (both inputs' blocks, the filters, the workgroups' work areas and
 * chunks of output nodes) is carved from one 128 MiB array in device memory
 * that the library holds on each GPU it runs on until the process ends; no
 * hipMalloc, hipFree or pinned host allocation is made.  Inputs whose blocks
 * do not fit beside the work areas (about 100 MiB at k = 7, about 85 MiB at
 * k = 11) are MWG_ERR_ARG.  The result equals mwg_multiply's bit for bit.
 * Calls on one GPU share the array: calls that overlap on one GPU (from
 * several threads or contexts) race and must not be made. */
int mwg_multiply_noalloc(mwg_context ctx, double prec, mwg_tree out, double coef, mwg_tree a, mwg_tree b,
                         int max_iter, int abs_prec);
/* out = sum_i coefs[i] inputs[i]: mrcpp::add(prec, out, FunctionTreeVector,
 * max_iter, abs_prec), which MRChem calls through
 * mrcpp::cplxfunc::linear_combination for the exchange sums
 * (src/qmoperators/two_electron/ExchangePotentialD1.cpp:317,366).  Each
 * output node's block is summed on the device from the inputs' blocks at the
 * node (below an input's finest node: its scaling part reconstructed from
 * there).  prec < 0: on the union of the input grids; prec >= 0: built from
 * the roots, a node split while the sum's wavelet norm fails the split check
 * (absolute with abs_prec, else relative to the sum's norm), for at most
 * max_iter levels (< 0: no limit).  out must differ from every input. */
int mwg_add(mwg_context ctx, double prec, mwg_tree out, int n, const double *coefs, const mwg_tree *inputs,
            int max_iter, int abs_prec);
/* tree = tree + coef * inp on the union of the two grids, in place:
 * FunctionTree::add(c, inp) (ExchangePotentialD1.cpp:367). */
int mwg_tree_add(mwg_context ctx, mwg_tree tree, double coef, mwg_tree inp);
/* Remove the children of every branch node whose children are all leaves
 * and whose wavelet norm is at most max(2 eps, prec * t * 2^{-split_fac
 * (scale + 1) / 2}) (t = 1 with abs_prec, else the tree's norm), bottom-up:
 * FunctionTree::crop(prec, splitFac, absPrec) (ExchangePotentialD1.cpp:323,371).
 * Host operation on the tree structure. */
int mwg_tree_crop(mwg_tree tree, double prec, double split_fac, int abs_prec);
/* Packed form of a tree: one contiguous little-endian buffer, the wire format
 * for sending a tree in one message (MRChem sends trees with
 * mrcpp::send_tree in chunks, src/parallel.cpp:325-336) and the file format
 * of mwg_tree_save (MRChem writes orbitals with saveOrbital,
 * src/qmfunctions/orbital_utils.cpp:397-401).  Layout: a 72-byte header
 * (magic "MWGTREE1", int32 version 1, kp1, basis type, root scale,
 * corner[3], boxes[3], double scaling factor, int64 node count n, double
 * squared norm), int32 depth[n] zero-padded to a multiple of 8 bytes, int64
 * l[n][3], double blocks[n][8 (k+1)^3], nodes parents first.  Unpacking
 * requires the same MRA (order, basis, root scale, corner, boxes, scaling
 * factor) and a complete tree. */
int mwg_tree_pack_size(mwg_tree tree, long long *bytes);
int mwg_tree_pack(mwg_tree tree, void *buffer, long long bytes);
int mwg_tree_unpack(mwg_tree tree, const void *buffer, long long bytes);
/* The packed form of mwg_tree_pack written into device memory of ctx's GPU
 * (dst: a device pointer to at least mwg_tree_pack_size bytes), for a
 * device-to-device send of a tree, e.g. over RCCL, in the chunks
 * mrcpp::send_tree sends (src/parallel.cpp:323-337): the header and keys are
 * uploaded and the blocks copied on the device (from a resident tree's
 * device copy, else uploaded once). */
int mwg_tree_pack_device(mwg_tree tree, mwg_context ctx, void *dst, long long bytes);
/* Replace the tree by a packed tree held in device memory of ctx's GPU (as
 * received from a device-to-device send).  The tree becomes resident on ctx
 * (mwg_tree_set_resident) and its blocks are gathered on the device; only
 * the header and keys come to the host. */
int mwg_tree_unpack_device(mwg_tree tree, mwg_context ctx, const void *src, long long bytes);
/* Device memory on ctx's GPU for packed trees (a send or receive buffer),
 * and a copy of it to host memory: for callers without an allocator of
 * their own on that device.  No counterpart in the reference. */
int mwg_device_buffer_create(mwg_context ctx, long long bytes, void **dev);
int mwg_device_buffer_destroy(mwg_context ctx, void *dev);
int mwg_device_buffer_read(mwg_context ctx, const void *dev, void *host, long long bytes);
int mwg_tree_save(mwg_tree tree, const char *path);
int mwg_tree_load(mwg_tree tree, const char *path);
/* Screening threshold of every node used by the last apply into this tree. */
int mwg_tree_thresholds(mwg_tree tree, double *thr);

/* mrcpp::apply(prec, out, oper, inp, max_iter, abs_prec): adaptive output
 * grid grown from the roots, compressed output with norm set. */
int mwg_apply(mwg_context ctx, double prec, mwg_tree out, mwg_operator op, mwg_tree inp, int max_iter,
              int abs_prec);
/* mrcpp::apply(prec, out, oper, inp, precTrees, max_iter, abs_prec). */
int mwg_apply_prec_trees(mwg_context ctx, double prec, mwg_tree out, mwg_operator op, mwg_tree inp,
                         const mwg_tree *prec_trees, int n_prec, int max_iter, int abs_prec);
/* One apply of a batch: mrcpp::apply(prec, out, op, inp[, prec_trees],
 * max_iter, abs_prec). */
typedef struct {
    mwg_tree out;
    mwg_operator op;
    mwg_tree inp;
    double prec;
    int max_iter;
    int abs_prec;
    const mwg_tree *prec_trees; /* may be NULL when n_prec == 0 */
    int n_prec;
} mwg_apply_desc;
/* Many applies at once, with the result of calling mwg_apply_prec_trees on
 * each entry: every refinement level of every apply still refining is
 * computed in the same launches.  This is the batch MRChem runs as sequential
 * loops: all pair densities of ExchangePotentialD1::setupInternal
 * (src/qmoperators/two_electron/ExchangePotentialD1.cpp:158-236) and all
 * orbitals of HelmholtzVector::apply (src/scf_solver/HelmholtzVector.cpp:141-159).
 * Output trees must be distinct and no output may be an input; operators and
 * inputs may be shared.  All trees and operators must share one MRA. */
int mwg_apply_batch(mwg_context ctx, int n, const mwg_apply_desc *desc);

/* Frozen-grid apply: out's current grid is kept, thr[g] (tree order) is the
 * screening threshold of node g; out receives the compressed result. */
int mwg_apply_fixed(mwg_context ctx, mwg_tree out, mwg_operator op, mwg_tree inp, const double *thr, int mode);

/* Statistics of the last apply on this context: [0] kernel milliseconds,
 * [1] output nodes computed, [2] input slots, [3] generated slots,
 * [4] host plan milliseconds, [5] kernel launches, [6] term-applications
 * that passed screening (each 6(k+1)^4 flops of algorithmic work),
 * wall milliseconds of [7] device plan creation, runs and downloads,
 * [8] the host split checks, [9] the post transforms (TopDown, BottomUp,
 * norms), [10] the whole apply, of the parts of [7]: [11] device plan
 * creation, [12] runs, and of the parts of [4]: [13] bands and slots,
 * [14] input pools on the device, [15] exact counts and segments (adaptive
 * applies; frozen-grid calls leave [11..15] 0). */
int mwg_last_apply_stats(mwg_context ctx, double *stats, int n);

/* Frozen-grid apply split into a resident device plan (bands, input slots,
 * operator tables and output blocks in HBM) and repeated executions on it.
 * mwg_apply_fixed = create + execute + finish + destroy.  No counterpart in
 * MRChem's calls; it serves repeated applies on one grid and measurement. */
int mwg_plan_create(mwg_context ctx, mwg_tree out, mwg_operator op, mwg_tree inp, const double *thr, int mode,
                    mwg_plan *plan);
/* Batched frozen-grid plan: n applies (out[i], op[i], inp[i], thresholds
 * thr[i]) resident together and run in the same launches. */
int mwg_plan_create_batch(mwg_context ctx, int n, const mwg_tree *out, const mwg_operator *op, const mwg_tree *inp,
                          const double *const *thr, int mode, mwg_plan *plan);
/* One execution of the per-node calculator on resident data, with the plan's
 * `mode` (tau from the context).  stats (n <= 5 values): [0] device ms,
 * [1] term-applications that passed screening, [2] of them on the
 * low-precision path (FP32 or FP16 split),
 * [3] device ms of the FP64-class launch, [4] of the low-precision launch
 * (MIXED, SPLIT16: the two run concurrently on two streams, [0] covers
 * both; FP64: [3] is the calculator launch). */
int mwg_plan_execute(mwg_plan plan, double *stats, int n);
/* One whole frozen-grid apply on resident data: the per-node calculator as
 * in mwg_plan_execute, then the TopDown(add) and BottomUp transforms and the
 * node norms of the output, all on the device and in place on the plan's
 * output blocks (what mrcpp::apply does after its calculator).  stats (n <= 6
 * values): [0]-[4] as mwg_plan_execute, [0] now covering the whole apply,
 * [5] device ms of the post transforms and norms.  No counterpart in
 * MRChem's calls; mwg_plan_finish then only downloads. */
int mwg_plan_apply(mwg_plan plan, double *stats, int n);
/* Autotuner of a frozen-grid plan.  Runs the plan once in FP64 as the
 * reference, then searches, one knob at a time from the current setting:
 * the FP64-class cut tau (MIXED, SPLIT16: terms with bound > tau x threshold
 * stay FP64), the single-rounding cut of the FP16 split class (SPLIT16), and
 * the order of the work queue (segment order, or the largest passing-term
 * count first).  Each configuration is timed over `reps` executions (device
 * time) and counts only if its largest output difference to the reference is
 * at most max_rel_err of the reference's largest value.  The fastest is kept
 * in the plan and cached in the context under (k + 1, operator terms, widest
 * band, mode): later plans with that key start from it, and the pick is
 * never slower than the starting setting when that one meets the bound.
 * report (n <= 9): [0] tau, [1] single-rounding ratio, [2] order, [3] its
 * device ms, [4] device ms of the starting setting, [5] its relative error,
 * [6] configurations timed, [7] speedup [4]/[3], [8] relative error of the
 * starting setting.  No counterpart in MRChem's calls. */
int mwg_plan_tune(mwg_plan plan, double max_rel_err, int reps, double *report, int n);
/* With on != 0, later mwg_plan_execute / mwg_plan_apply calls replay one
 * hipGraph of the whole step (output and counters cleared by kernels, the
 * calculator launch or launches, and for mwg_plan_apply the post transforms
 * and norms), captured on the plan after the first eager run of that step
 * and kept until the plan is destroyed; a replay reports the whole device
 * time only (stats [3]-[5] are 0).  Default off.  No counterpart in MRChem's
 * calls; it serves repeated applies on one grid. */
int mwg_plan_use_graph(mwg_plan plan, int on);
/* Copy the result into the output tree(s) and run the TopDown/BottomUp
 * transforms (after mwg_plan_apply: download the transformed blocks).  MWG_ERR_ARG when an output tree's grid is no longer the one
 * the plan was made on. */
int mwg_plan_finish(mwg_plan plan);
/* Raw non-standard blocks of the last execution (n_out x 8(k+1)^3, tree
 * order, applies of a batch one after the other), before the TopDown/BottomUp
 * transforms.  capacity: doubles ns holds; MWG_ERR_ARG when it is less than
 * n_out x 8(k+1)^3 (mwg_plan_info [0] gives n_out), or when the last
 * execution was mwg_plan_apply (its blocks are transformed in place). */
int mwg_plan_download_ns(mwg_plan plan, double *ns, long long capacity);
/* [0] output nodes, [1] input slots, [2] generated slots, [3] host plan ms,
 * [4] term-applications counted on the host with the screening rule,
 * [5] work segments (entries of the FP64 work queue), [6] applies,
 * [7] widest band half-width (translations |l| <= W per direction),
 * [8] FP64-class cut tau, [9] single-rounding ratio, [10] queue order. */
int mwg_plan_info(mwg_plan plan, double *info, int n);
int mwg_plan_destroy(mwg_plan plan);

#ifdef __cplusplus
}
#endif

#endif /* MWGREEN_H */
