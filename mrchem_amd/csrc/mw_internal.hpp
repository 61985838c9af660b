// Host-side multiwavelet machinery shared by the C ABI and the HIP apply.
//
// Everything here is FP64 C++17.  It restates the published MRCPP design
// (Alpert multiwavelets, Beylkin-Coifman-Rokhlin non-standard operator form,
// separated Gaussian expansions of 1/r and exp(-mu r)/r) that MRChem calls
// through mrcpp::apply (src/qmoperators/two_electron/CoulombPotential.cpp:131,
// src/qmoperators/two_electron/ExchangePotential.cpp:200,
// src/scf_solver/HelmholtzVector.cpp:149).
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

namespace mwg {

// ------------------------------------------------------------------- threads
// Threads of the library's parallel loops: MWG_NUM_THREADS when it is a
// positive count; else half the CPUs the process may use (affinity mask and
// cgroup CPU quota), at least 1 and at most OpenMP's own count.  The loops
// are short (milliseconds) and each waits for its slowest thread, so a team
// as large as the CPU count loses whenever the GPU runtime's or the caller's
// threads take a CPU: on a 16-CPU host a 2688-node adaptive apply took
// 174 ms with 16 threads, 107 ms with 12 and 77 ms with 8.
int team_size();

// ---------------------------------------------------------------- math utils
// n-point Gauss-Legendre rule mapped to [0,1] (nodes ascending, weights sum to 1).
void gauss_legendre01(int n, double *x, double *w);
// Normalised Legendre functions on [0,1]: P~_p(x) = sqrt(2p+1) P_p(2x-1), p < n.
void legendre01(int n, double x, double *out);
// Moments M_p = int_0^1 P~_p(v) exp(-A (v-c)^2) dv for p < n, accurate to ~1e-16
// relative to int_0^1 |P~_p| exp(...) for any A >= 0 and real c.
void gauss_moments01(int n, double A, double c, double *out);

// ------------------------------------------------------------------ basis
enum BasisType { BASIS_INTERPOLATING = 0, BASIS_LEGENDRE = 1 };

struct Basis {
    int k = 0;    // polynomial order
    int kp1 = 0;  // functions per direction
    int type = BASIS_INTERPOLATING;
    // phi_i(x) = sum_p L2B[i*kp1+p] P~_p(x)  (orthogonal matrix)
    std::vector<double> L2B;
    // Two-scale filter, 2kp1 x 2kp1 row-major: rows [H0 H1 ; G0 G1].
    // [s^n ; d^n] = F [s^{n+1}_{2l} ; s^{n+1}_{2l+1}]   (F orthogonal)
    std::vector<double> F;
    // Cross-correlation: psi_ij(u) = int phi_i(x) phi_j(x-u) dx restricted to
    // u in [0,1] equals sum_p C[(i*kp1+j)*2kp1 + p] P~_p(u), p < 2kp1.
    std::vector<double> C;
    void eval(double x, double *out) const;  // phi_i(x), x in [0,1]
};
Basis make_basis(int k, int type);

// ------------------------------------------------------------------- MRA
struct MRA {
    Basis basis;
    int n0 = 0;                // root scale
    int corner[3] = {-1, -1, -1};
    int boxes[3] = {2, 2, 2};
    double sfac = 1.0;         // physical side of a scale-0 box
    int max_depth = 20;        // deepest depth (scale n0 + max_depth)
    int kp1() const { return basis.kp1; }
    int kp1_3() const { return basis.kp1 * basis.kp1 * basis.kp1; }
    int block() const { return 8 * kp1_3(); }
    double box_size(int depth) const;        // physical side at depth
    double world_lo(int d) const;            // physical lower bound along d
    double world_hi(int d) const;
    long long tlo(int depth, int d) const { return (long long)corner[d] << depth; }
    long long thi(int depth, int d) const { return (long long)(corner[d] + boxes[d]) << depth; }  // exclusive
    bool in_world(int depth, const long long *l) const;
};

// --------------------------------------------------------- kernel fitting
struct GaussTerm {
    double alpha;  // exponent (physical units)
    double beta;   // 3D coefficient
};
// Separated expansion K(r) ~= sum_m beta_m exp(-alpha_m r^2) with
// |K_fit(r) - K(r)| * r <= eps on [r_min, r_max].  mu == 0 gives 1/r,
// mu > 0 gives exp(-mu r)/r.
std::vector<GaussTerm> fit_kernel(double mu, double eps, double r_min, double r_max,
                                  double *achieved_err = nullptr);
double kernel_fit_error(const std::vector<GaussTerm> &t, double mu, double r_min, double r_max,
                        int npts);

// ----------------------------------------------------- operator tables
// One separated term: per depth, the four non-standard 1D components
// c = 2*gt + ft (0: s<-s, 1: s<-d, 2: d<-s, 3: d<-d) for translations
// l in [-L, L], each a kp1 x kp1 block O[i][j] = <out_i | K_m | in_j>.
struct OperDepth {
    int L = -1;                  // -1: nothing stored at this depth
    std::vector<double> blocks;  // [c][l+L][i][j]
    std::vector<double> norms;   // [c][l+L] Frobenius norms
};
struct OperTerm {
    double alpha = 0, beta = 0, coef1d = 0;
    std::vector<OperDepth> depth;
};
struct Operator {
    const MRA *mra = nullptr;
    double mu = 0.0;          // 0 for Poisson
    double build_prec = 0.0;
    double r_min = 0, r_max = 0, fit_err = 0;
    int max_depth = 0;
    std::vector<OperTerm> terms;
    int size() const { return (int)terms.size(); }
};
// Build the operator for kernel exp(-mu r)/r (mu = 0: 1/r).
void build_operator(Operator &op, const MRA &mra, double mu, double build_prec, int max_depth);
// Tables for explicitly given Gaussian terms (tests and custom kernels).
// Blocks below tau relative to the term's largest block at that depth are not
// stored (see build_term_depth).
void build_operator_terms(Operator &op, const MRA &mra, const std::vector<GaussTerm> &g, int max_depth,
                          double tau = 1e-14);
// Direct T-block at depth for translation l (used by tests and construction).
void t_block(const MRA &mra, double alpha, double coef1d, int depth, long long l, double *out);

// ----------------------------------------------------------------- trees
struct NodeKey {
    int depth;
    long long l[3];
};
uint64_t morton3(uint64_t x, uint64_t y, uint64_t z);  // 21 bits per dimension
struct KeyHash {
    size_t operator()(const std::pair<int, uint64_t> &k) const {
        return std::hash<uint64_t>()(k.second * 0x9E3779B97F4A7C15ULL ^ (uint64_t)k.first);
    }
};

// Allocator for large host arrays (tree coefficients, input pools).
// Requests of at least 4 MiB are 2 MiB aligned and marked for transparent
// huge pages (MADV_HUGEPAGE), so a multi-GB tree is first touched in 2 MiB
// pages instead of 4 KiB ones (a 13 GB tree: 0.43 s instead of 1.17 s of
// page faults on 8 threads); smaller requests come from operator new.
// INIT = false leaves new elements uninitialised (default-init), for
// buffers written in full right after allocation: their pages are then
// first touched by the (parallel) writer instead of a serial zero fill.
constexpr size_t kBigAllocMin = (size_t)4 << 20;
constexpr size_t kHugePage = (size_t)2 << 20;
void *big_alloc(size_t bytes);
void big_free(void *p, size_t bytes);

template <typename T, bool INIT>
struct BigAlloc {
    using value_type = T;
    template <typename U>
    struct rebind {
        using other = BigAlloc<U, INIT>;
    };
    BigAlloc() noexcept = default;
    template <typename U>
    BigAlloc(const BigAlloc<U, INIT> &) noexcept {}
    T *allocate(size_t n) { return static_cast<T *>(big_alloc(n * sizeof(T))); }
    void deallocate(T *p, size_t n) noexcept { big_free(p, n * sizeof(T)); }
    template <typename U, typename... Args>
    void construct(U *p, Args &&...args) {
        if constexpr (!INIT && sizeof...(Args) == 0)
            ::new ((void *)p) U;
        else
            ::new ((void *)p) U(std::forward<Args>(args)...);
    }
    template <typename U>
    bool operator==(const BigAlloc<U, INIT> &) const noexcept {
        return true;
    }
    template <typename U>
    bool operator!=(const BigAlloc<U, INIT> &) const noexcept {
        return false;
    }
};
template <typename T>
using UninitAlloc = BigAlloc<T, false>;
// Tree coefficients: a resize without a value leaves new elements
// uninitialised (add_node fills its block with zeros explicitly).
using CoefVec = std::vector<double, BigAlloc<double, false>>;

struct Tree {
    const MRA *mra = nullptr;
    std::vector<NodeKey> keys;
    std::vector<int> parent;   // -1 for roots
    std::vector<int> child0;   // first of 8 contiguous children, -1 for leaves
    CoefVec coefs;             // nodes x block; compressed [s, d1..d7] at every node
    // structure only: add_node leaves coefs alone (an apply grows its output
    // grid this way and sizes coefs once, for the blocks it then writes)
    bool structure_only = false;
    std::unordered_map<std::pair<int, uint64_t>, int, KeyHash> index;
    double sqnorm = 0.0;

    int size() const { return (int)keys.size(); }
    uint64_t morton(const NodeKey &k) const;
    int find(int depth, const long long *l) const;
    int add_node(const NodeKey &k, int par);   // coefs zero-initialised
    void make_roots();
    int split(int node);                       // create 8 children, returns child0
    double *node(int i) { return coefs.data() + (size_t)i * mra->block(); }
    const double *node(int i) const { return coefs.data() + (size_t)i * mra->block(); }
    void clear();
    int max_depth() const;
    int n_leaves() const;
};

// 3D two-scale transforms of one node block (8 components of kp1^3, x fastest).
// compress: in = 8 child scaling blocks (child t = bx + 2 by + 4 bz),
//           out = [s, d1..d7] (component bit d set = wavelet along d).
void mw_compress(const Basis &b, const double *in, double *out);
void mw_reconstruct(const Basis &b, const double *in, double *out);
// Scaling block of child t alone (t = bx + 2 by + 4 bz); `in` is a full
// compressed block, or only its scaling part when s_only.
void mw_child_scaling(const Basis &b, const double *in, bool s_only, int t, double *out);
void component_norms(const MRA &mra, const double *blk, double *norms8);
double tree_square_norm(const Tree &t);
double tree_dot(const Tree &a, const Tree &b);
// bottom-up: recompute every branch node's [s,d] from its children's s parts
void tree_compress(Tree &t);

// Gaussian sums: f(r) = sum_i c_i (a_i/pi)^{3/2} exp(-a_i |r - R_i|^2)
struct GaussSum {
    std::vector<double> coef, exp, pos;  // pos: 3 per term
    int size() const { return (int)coef.size(); }
    double eval(const double *r) const;
    double square_norm() const;
};
// Scaling coefficients of f on node (depth, l), kp1^3 block (exact integrals).
void project_scaling(const MRA &mra, const GaussSum &f, int depth, const long long *l, double *out);
// Adaptive projection: split while ||d|| > prec * tnorm * 2^{-(n+1)/2}.
void project_adaptive(Tree &t, const GaussSum &f, double prec, bool abs_prec, int max_depth);
// Evaluate a compressed tree at a physical point.
double tree_eval(const Tree &t, const double *r);

// Split test used by projection and by the adaptive apply.
bool split_check(const MRA &mra, int depth, double wnorm, double prec, double tnorm);

// ----------------------------------------------------------- apply plan
// Everything the per-node calculator needs, laid out flat: the output nodes
// (one refinement level or a whole fixed grid), the input nodes of their
// bands (real or generated by reconstruction), the dense band table, and the
// screening thresholds.  Built on the host; consumed by the HIP kernels.
struct ApplyPlan {
    int kp1 = 0, ndepth = 0, M = 0;
    // output nodes (in the order of out_idx)
    std::vector<int> out_node;      // node index in output tree
    std::vector<int> out_depth;
    std::vector<double> out_thr;    // screening threshold per output node
    std::vector<long long> band_off;  // offset of this node's band table entries
    std::vector<int> band_w;        // band half-width W (band has (2W+1)^3 entries)
    std::vector<int> band;          // input slot or -1, order dz,dy,dx with dx fastest
    // input slots (real or generated nodes)
    std::vector<long long> slot_off;  // offset of the slot's block in in_coefs
    std::vector<double, UninitAlloc<double>> in_coefs;  // pool: full blocks (real) or scaling part (generated)
    std::vector<double> in_norms;   // slots x 8 (wavelet norms 0 for generated slots)
    std::vector<int> in_depth;
    // screening-derived bandwidths: w[(d*M + m)*4 + c]
    std::vector<int> bw;
    std::vector<int> active_off, active_terms;  // per depth, terms with any w >= 0
    long long n_generated = 0;
    // Work segments of the output component pairs (gt, gt | 4), gt < 4:
    // job = 8*out + gt; a segment covers the (band position p, input
    // component ft) pairs 8 p + ft in [pseg_p0, pseg_p1) of one job (whole
    // band positions, except in small launches, where a band position with
    // more terms than a segment is cut by ft).  Jobs are cut so that no
    // segment holds many more than
    // `seg_target` term-applications of the two components (counted exactly
    // here with the kernels' screening rule); jobs without any are left out.
    // The kernels share the x and y passes of a pair.  Per segment: passing
    // terms and the smallest / largest screening bound among them (for
    // choosing the segments of each precision class).
    std::vector<int> pseg_job, pseg_p0, pseg_p1, pseg_apps;
    std::vector<double> pseg_bmin, pseg_bmax;
    // Fine segments: single-component jobs cut at about seg_target / 8
    // passing terms, for the launch of the FP64 class of the mixed and split
    // modes, whose terms crowd into few segments (coarse output nodes).
    std::vector<int> fseg_job, fseg_p0, fseg_p1, fseg_apps;
    std::vector<double> fseg_bmin, fseg_bmax;
    long long term_apps = 0;   // exact count of term-applications in the plan
    double thr_floor = 0.0;    // smallest positive threshold of the plan
    // A pool made on the device (build_plan with device_pool): in_coefs
    // stays empty; slot_src[s] is the slot's input-tree node (>= 0) or
    // generated node -(g + 1); generated node g is child gen_child[g] of
    // generated node gen_parent[g] (>= 0) or of input node -(gen_parent[g]
    // + 1), in order of depth, [gen_level_off[i], gen_level_off[i + 1]) one
    // depth; slot_off[s] is the slot's offset in the pool as before.
    bool device_pool = false;
    std::vector<int> slot_src, gen_parent, gen_child, gen_level_off;
    long long pool_values = 0;  // doubles of the pool
};

struct ApplyStats {
    double term_apps = 0;      // term-applications passing screening
    double candidates = 0;     // (f, ft, gt, m) candidates examined
    double out_nodes = 0;
    double in_slots = 0;
    double generated = 0;
};

// Build the plan for output nodes `nodes` of `out` with screening parameter
// thr_param (thr_g = thr_param * prec_fac[g]).
// With device_pool the plan stops after the slots and their sources: the
// pool and in_norms are made on the device (dev_pool_fill), then
// finish_plan_counts adds the exact counts and the work segments.
// in_cnorm8: the input tree's 8 component norms per node (tree order), when
// the caller has them; otherwise computed here.
void build_plan(ApplyPlan &plan, const Operator &op, const Tree &inp, const Tree &out,
                const std::vector<int> &nodes, const std::vector<double> &thr,
                int max_band_w, int seg_target = 256, bool device_pool = false,
                const double *in_cnorm8 = nullptr);
// fine_segments: also cut the fine single-component segments (the FP64 class
// of the mixed and split modes; plain FP64 runs over the pair segments only).
// n_share: the number of plans of about this size launched together (the
// applies of a batch level), which decides whether the launch is small.
void finish_plan_counts(ApplyPlan &plan, const Operator &op, int seg_target = 256, bool fine_segments = true,
                        int n_share = 1);
double oper_norm(const Operator &op, int m, int d, int c, long long l);

// Operator tables flattened for the device: per (m, d) the half-width L, the
// offset of block (c=0, l=-L) in `blocks` (doubles) and of its norm in `norms`.
struct OperFlat {
    int M = 0, ndepth = 0, kp1 = 0;
    std::vector<int> L;             // [m*ndepth + d]
    std::vector<long long> boff, noff;
    std::vector<double> blocks, norms;
    // > 0: the blocks were built on a device and live only there
    // (this many values); blocks is then empty
    long long device_blocks = 0;
};
void flatten_operator(const Operator &op, OperFlat &f);
// Storage floor parameter of an operator built to build_prec.
double operator_tau(double build_prec);
// Operator header without tables: world extent, kernel fit and the terms
// (alpha, beta, coef1d), each with max_depth + 1 empty depths.
void prepare_operator(Operator &op, const MRA &mra, double mu, double build_prec, int max_depth);
// Translation scan of one (term, depth): false when the depth stores nothing;
// otherwise the scan bound Lscan and the position `peak` after which three
// quiet translations end the scan.
bool term_depth_scan(const MRA &mra, double alpha, int d, double tau, long long &Lscan, double &peak);
// Gauss-Legendre rule of n points on [0, 1] (cached).
void quad_rule01(int n, const double **x, const double **w);
// Pre-screen table of an apply plan for the device screening: for each
// output depth d in [d0, d0 + nd), component combination cmb = cx | cy << 2 |
// cz << 4 and translation |l| <= w (the plan's widest band), the largest
// ((N_x N_y) N_z) over the depth's active terms.  The tables of all plans of
// a batch are built in one parallel pass over (plan, depth, cmb).
void build_prescreens(const std::vector<const ApplyPlan *> &pls, const std::vector<const OperFlat *> &fs,
                      std::vector<std::vector<double>> &mts, std::vector<std::pair<int, int>> &shapes);

}  // namespace mwg
