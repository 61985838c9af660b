// Pointwise product of two multiwavelet functions, node by node (gfx950).
//
// mrcpp::multiply(prec, out, c, a, b, maxIter, absPrec), reached from MRChem
// through mrcpp::cplxfunc::multiply for the pair densities rho_ij = phi_i phi_j
// (src/qmoperators/two_electron/ExchangePotential.cpp:175) and the exchange
// products phi_k V_ij (ExchangePotential.cpp:213,223).  In the interpolating
// basis the scaling coefficients of a node are the function's values at the
// node's Gauss-Legendre points, scaled: f(x_ijk) = s_ijk h^{-3/2} /
// sqrt(w_i w_j w_k) for a box of side h.  For one output node at depth n the
// device forms both inputs' scaling blocks at its 8 children (reconstructing
// from each input's deepest node on the path, wavelet parts zero below it),
// multiplies them point by point, c s^a s^b h_{n+1}^{-3/2} / sqrt(w_i w_j w_k),
// and compresses the 8 children into the node's [s, d1..d7] block.  The host
// keeps the grid and the refinement loop (WaveletAdaptor's split rule).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>

#include "mw_device.hpp"

namespace mwg {

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);                       \
            return -2;                                                                     \
        }                                                                                  \
    } while (0)

namespace {

constexpr int kMulThreads = 256;
constexpr int kMulBlocks = 1024;  // resident workgroups, each with its own work area

// Index of the element at position i along dim of the line (a, c) in an
// n^3 block (x fastest): the two other indices are a (the lower direction)
// and c (the higher one), as in the host transforms (mw_tree.cpp).
__device__ __forceinline__ int line_index(int n, int dim, int a, int c, int i) {
    return dim == 0 ? a * n + c * n * n + i : (dim == 1 ? a + c * n * n + i * n : a + c * n + i * n * n);
}

// One direction of the two-scale transform on all 4 component pairs
// (t, t | bit) of an 8-component block, src -> dst: compress (M = F) or
// reconstruct (M = F^T): [u_lo ; u_hi] = M [v_lo ; v_hi] along dim.
__device__ void pair_pass(int n, const double *F, const double *src, double *dst, int dim, bool transpose, int tid) {
    const int n2 = 2 * n, n3 = n * n * n, bit = 1 << dim;
    for (int e = tid; e < 8 * n3; e += kMulThreads) {
        const int t = e / n3, rem = e % n3;
        // rem = line_index(dim, a, c, i): recover (a, c, i)
        int i, a, c;
        if (dim == 0) {
            i = rem % n;
            a = (rem / n) % n;
            c = rem / (n * n);
        } else if (dim == 1) {
            a = rem % n;
            i = (rem / n) % n;
            c = rem / (n * n);
        } else {
            a = rem % n;
            c = (rem / n) % n;
            i = rem / (n * n);
        }
        const int t0 = t & ~bit, t1 = t0 | bit;
        const int r = (t & bit) ? n + i : i;
        const double *lo = src + (size_t)t0 * n3, *hi = src + (size_t)t1 * n3;
        double s = 0.0;
        for (int j = 0; j < n; j++) {
            const int q = line_index(n, dim, a, c, j);
            const double mlo = transpose ? F[j * n2 + r] : F[r * n2 + j];
            const double mhi = transpose ? F[(n + j) * n2 + r] : F[r * n2 + n + j];
            s += mlo * lo[q] + mhi * hi[q];
        }
        dst[e] = s;
    }
}

// One direction of the inverse step towards child half `bit` for nb block
// pairs: dst[k] = sum_j H_bit[j][i] lo_k[j] + G_bit[j][i] hi_k[j] along dim,
// with lo_k = src + lo_off[k] blocks, hi_k likewise (hi_off < 0: none).
__device__ void child_pass(int n, const double *F, const double *src, const int *lo_blk, const int *hi_blk, int nb,
                           double *dst, int dim, int bit, int tid, int nth = kMulThreads) {
    const int n2 = 2 * n, n3 = n * n * n;
    for (int e = tid; e < nb * n3; e += nth) {
        const int k = e / n3, rem = e % n3;
        int i, a, c;
        if (dim == 0) {
            i = rem % n;
            a = (rem / n) % n;
            c = rem / (n * n);
        } else if (dim == 1) {
            a = rem % n;
            i = (rem / n) % n;
            c = rem / (n * n);
        } else {
            a = rem % n;
            c = (rem / n) % n;
            i = rem / (n * n);
        }
        const double *lo = src + (size_t)lo_blk[k] * n3;
        const double *hi = hi_blk[k] >= 0 ? src + (size_t)hi_blk[k] * n3 : nullptr;
        double s = 0.0;
        for (int j = 0; j < n; j++) {
            const int q = line_index(n, dim, a, c, j);
            s += F[j * n2 + bit * n + i] * lo[q];
            if (hi) s += F[(n + j) * n2 + bit * n + i] * hi[q];
        }
        dst[(size_t)k * n3 + rem] = s;
    }
}

// Scaling block of the node `dep` >= 1 levels below the source node block
// `src` (8 components) along `path` (3 bits per level, first level lowest)
// into X[0, n^3); X holds 4 and Y 2 blocks of work area.
__device__ void descend(int n, const double *F, const double *src, int dep, unsigned long long path, double *X,
                        double *Y, int tid) {
    const int n3 = n * n * n;
    // first level: the full block (scaling and wavelet parts) towards child t
    const int t = path & 7;
    {
        const int lo[4] = {0, 1, 2, 3}, hi[4] = {4, 5, 6, 7};
        child_pass(n, F, src, lo, hi, 4, X, 2, (t >> 2) & 1, tid);
    }
    __syncthreads();
    {
        const int lo[2] = {0, 1}, hi[2] = {2, 3};
        child_pass(n, F, X, lo, hi, 2, Y, 1, (t >> 1) & 1, tid);
    }
    __syncthreads();
    {
        const int lo[1] = {0}, hi[1] = {1};
        child_pass(n, F, Y, lo, hi, 1, X, 0, t & 1, tid);
    }
    __syncthreads();
    // further levels: scaling part only
    for (int lv = 1; lv < dep; lv++) {
        const int tl = (path >> (3 * lv)) & 7;
        const int lo[1] = {0}, none[1] = {-1};
        child_pass(n, F, X, lo, none, 1, Y, 2, (tl >> 2) & 1, tid);
        __syncthreads();
        child_pass(n, F, Y, lo, none, 1, Y + n3, 1, (tl >> 1) & 1, tid);
        __syncthreads();
        child_pass(n, F, Y + n3, lo, none, 1, X, 0, tl & 1, tid);
        __syncthreads();
    }
}

// Scaling blocks of the 8 children of the work node, from the source node
// block `src` (8 components) `dep` levels above it along `path`, into X
// (Y: work area of 8 blocks).
__device__ void children_of(int n, const double *F, const double *src, int dep, unsigned path, double *X, double *Y,
                            int tid) {
    const int n3 = n * n * n;
    if (dep == 0) {
        for (int e = tid; e < 8 * n3; e += kMulThreads) Y[e] = src[e];
        __syncthreads();
    } else {
        descend(n, F, src, dep, path, X, Y, tid);
        // the work node's block: scaling part X[0], wavelet parts zero
        for (int e = tid; e < 8 * n3; e += kMulThreads) Y[e] = e < n3 ? X[e] : 0.0;
        __syncthreads();
    }
    // reconstruct Y -> 8 children (directions z, y, x), ending in X
    pair_pass(n, F, Y, X, 2, true, tid);
    __syncthreads();
    pair_pass(n, F, X, Y, 1, true, tid);
    __syncthreads();
    pair_pass(n, F, Y, X, 0, true, tid);
    __syncthreads();
}

}  // namespace

__global__ __launch_bounds__(kMulThreads) void k_mul_nodes(int n, const MulNode *nodes, int n_nodes,
                                                           const double *pool_a, const double *pool_b,
                                                           const double *F, const double *wsc, double coef,
                                                           double *ws, double *out) {
    const int tid = threadIdx.x;
    const int n3 = n * n * n;
    double *X = ws + (size_t)blockIdx.x * 24 * n3, *Y = X + 8 * n3, *P = Y + 8 * n3;
    for (int w = blockIdx.x; w < n_nodes; w += gridDim.x) {
        const MulNode nd = nodes[w];
        children_of(n, F, pool_a + nd.src_a, nd.dep_a, nd.path_a, X, Y, tid);
        for (int e = tid; e < 8 * n3; e += kMulThreads) P[e] = X[e];
        __syncthreads();
        children_of(n, F, pool_b + nd.src_b, nd.dep_b, nd.path_b, X, Y, tid);
        // values product at the children's points, back to coefficients
        for (int e = tid; e < 8 * n3; e += kMulThreads) P[e] = coef * P[e] * X[e] * nd.hfac * wsc[e % n3];
        __syncthreads();
        // compress the 8 children (directions x, y, z) into [s, d1..d7]
        pair_pass(n, F, P, X, 0, false, tid);
        __syncthreads();
        pair_pass(n, F, X, Y, 1, false, tid);
        __syncthreads();
        double *dst = out + (size_t)w * 8 * n3;
        pair_pass(n, F, Y, dst, 2, false, tid);
        __syncthreads();
    }
}

// Sums: out block of node w = sum_i coefs[i] x (input i's block at the node,
// or, below the input's finest node on the path, the scaling part
// reconstructed from that node with zero wavelet parts).  src[w * n_in + i]
// locates input i's source node (off < 0: the input has no node on the path
// and contributes nothing).  Each workgroup owns its output blocks, so the
// sums need no atomics; every thread adds into the same elements for every
// input.
__global__ __launch_bounds__(kMulThreads) void k_add_nodes(int n, const AddSrc *src, int n_in, int n_nodes,
                                                           const double *pool, const double *coefs, const double *F,
                                                           double *ws, double *out) {
    const int tid = threadIdx.x;
    const int n3 = n * n * n;
    double *X = ws + (size_t)blockIdx.x * 16 * n3, *Y = X + 8 * n3;
    for (int w = blockIdx.x; w < n_nodes; w += gridDim.x) {
        double *dst = out + (size_t)w * 8 * n3;
        for (int e = tid; e < 8 * n3; e += kMulThreads) dst[e] = 0.0;
        for (int i = 0; i < n_in; i++) {
            const AddSrc sr = src[(size_t)w * n_in + i];
            if (sr.off < 0) continue;
            const double c = coefs[i];
            if (sr.dep == 0) {
                const double *b = pool + sr.off;
                for (int e = tid; e < 8 * n3; e += kMulThreads) dst[e] += c * b[e];
            } else {
                descend(n, F, pool + sr.off, sr.dep, sr.path, X, Y, tid);
                for (int e = tid; e < n3; e += kMulThreads) dst[e] += c * X[e];
                __syncthreads();  // X is rewritten for the next input
            }
        }
    }
}

// The generated children of one parent in an apply plan's pool: out[t] is
// the scaling block of child t (null when child t is not generated), the
// parent an input-tree node (full: its 8 components) or a generated node
// (its scaling block only).
struct GenGroup {
    const double *src;
    double *out[8];
    int full, mask;
};

// Threads of a k_gen_fill workgroup: its LDS (6 n^3 doubles, 83 KB at
// n = 12) allows one or two workgroups per compute unit, so a workgroup
// brings 16 waves (64 Helmholtz applies at k = 11, adaptive on resident
// trees: 1099 -> 1023 ms against 256 threads).
constexpr int kGenThreads = 1024;

// Generated nodes of one depth, grouped by parent (their parents are input
// nodes or generated nodes of the depth above): one workgroup per parent
// runs the z pass once per z half of the children it has, the y pass once
// per (z, y) quarter and the x pass per child, through LDS (X: 4 blocks,
// Y: 2); each output gets the operations it gets taken alone.
__global__ __launch_bounds__(kGenThreads) void k_gen_fill(int n, const double *F, const GenGroup *groups,
                                                          int n_groups) {
    extern __shared__ double lds[];  // 6 n^3
    const int tid = threadIdx.x, n3 = n * n * n;
    double *X = lds, *Y = lds + 4 * n3;
    for (int gi = blockIdx.x; gi < n_groups; gi += gridDim.x) {
        const GenGroup &gr = groups[gi];
        const int mask = gr.mask, full = gr.full;
        const double *src = gr.src;
        for (int zb = 0; zb < 2; zb++) {
            if (!((mask >> (4 * zb)) & 0xf)) continue;
            if (full) {
                const int lo[4] = {0, 1, 2, 3}, hi[4] = {4, 5, 6, 7};
                child_pass(n, F, src, lo, hi, 4, X, 2, zb, tid, kGenThreads);
            } else {
                const int lo[1] = {0}, none[1] = {-1};
                child_pass(n, F, src, lo, none, 1, X, 2, zb, tid, kGenThreads);
            }
            __syncthreads();
            for (int yb = 0; yb < 2; yb++) {
                if (!((mask >> (4 * zb + 2 * yb)) & 0x3)) continue;
                if (full) {
                    const int lo2[2] = {0, 1}, hi2[2] = {2, 3};
                    child_pass(n, F, X, lo2, hi2, 2, Y, 1, yb, tid, kGenThreads);
                } else {
                    const int lo[1] = {0}, none[1] = {-1};
                    child_pass(n, F, X, lo, none, 1, Y, 1, yb, tid, kGenThreads);
                }
                __syncthreads();
                for (int xb = 0; xb < 2; xb++) {
                    const int t = xb | (yb << 1) | (zb << 2);
                    if (!((mask >> t) & 1)) continue;
                    if (full) {
                        const int lo1[1] = {0}, hi1[1] = {1};
                        child_pass(n, F, Y, lo1, hi1, 1, gr.out[t], 0, xb, tid, kGenThreads);
                    } else {
                        const int lo[1] = {0}, none[1] = {-1};
                        child_pass(n, F, Y, lo, none, 1, gr.out[t], 0, xb, tid, kGenThreads);
                    }
                }
                __syncthreads();
            }
        }
    }
}

// One slot of an apply plan's pool: len doubles from src (an input node's
// full block, or a generated node's scaling block) to dst, and the slot's 8
// component norms to norms (wavelet norms 0 for a generated slot).
struct SlotItem {
    const double *src;
    double *dst, *norms;
    int len;
};

// The copies: one wave per slot.
__global__ __launch_bounds__(64) void k_pool_fill(const SlotItem *items, int n_items) {
    const int lane = threadIdx.x;
    for (int s = blockIdx.x; s < n_items; s += gridDim.x) {
        const SlotItem it = items[s];
        for (int e = lane; e < it.len; e += 64) it.dst[e] = it.src[e];
    }
}

// The norms: one lane per (slot, component), each summed in index order.
__global__ __launch_bounds__(256) void k_slot_norms(int n, const SlotItem *items, int n_items) {
    const int n3 = n * n * n;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= 8LL * n_items) return;
    const int s = (int)(i >> 3), comp = (int)(i & 7);
    const SlotItem it = items[s];
    double q = 0.0;
    if (comp * n3 < it.len) {
        const double *x = it.src + (size_t)comp * n3;
        for (int e = 0; e < n3; e++) q += x[e] * x[e];
    }
    it.norms[comp] = sqrt(q);
}

// Squared norms of each node's scaling part and wavelet parts: one
// workgroup of four waves per node; the four wave sums are added in wave
// order, so the result does not depend on timing.
__global__ __launch_bounds__(256) void k_node_norms(const double *blocks, long long n, int n3, double *s2, double *w2) {
    __shared__ double part[2][4];
    const long long b = blockIdx.x;
    if (b >= n) return;
    const double *x = blocks + b * 8 * n3;
    const int tid = threadIdx.x, wave = tid >> 6;
    double s = 0.0, w = 0.0;
    for (int e = tid; e < n3; e += 256) s += x[e] * x[e];
    for (int e = n3 + tid; e < 8 * n3; e += 256) w += x[e] * x[e];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        w += __shfl_xor(w, o);
    }
    if ((tid & 63) == 0) {
        part[0][wave] = s;
        part[1][wave] = w;
    }
    __syncthreads();
    if (tid == 0) {
        s2[b] = (part[0][0] + part[0][1]) + (part[0][2] + part[0][3]);
        w2[b] = (part[1][0] + part[1][1]) + (part[1][2] + part[1][3]);
    }
}

// The 8 component norms of each node (scaling part, then the 7 wavelet
// parts): one wave per node.
__global__ __launch_bounds__(64) void k_component_norms(const double *blocks, long long n, int n3, double *out8) {
    const long long b = blockIdx.x;
    if (b >= n) return;
    const double *x = blocks + b * 8 * n3;
    for (int t = 0; t < 8; t++) {
        double s = 0.0;
        for (int e = threadIdx.x; e < n3; e += 64) s += x[t * n3 + e] * x[t * n3 + e];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (threadIdx.x == 0) out8[b * 8 + t] = sqrt(s);
    }
}

// Post transforms (TopDown(add) and BottomUp of the apply's output tree).
// One workgroup per branch node: its block (TopDown) or its 8 children's
// scaling blocks (BottomUp) are staged in LDS, 8 N^3 doubles, and each of
// the three directions is one in-place pass over the 4 N^2 lines of the
// block: a thread takes a whole line (the N values of component t0 and the
// N of t1 = t0 | bit along the direction, bit = 1 << dim), multiplies the
// 2N-vector by the 2N x 2N filter (F for compress, F^T for reconstruct) in
// registers and writes it back.  Every lane of a wave is at the same filter
// element at the same time: the filter sits in LDS with column j of the
// pass's matrix contiguous (MT[j][r] = M[r][j]), its reads are broadcasts,
// and the 2N outputs of a line accumulate side by side (2N independent
// chains) while the inputs v[j] are taken in turn.
// Threads per workgroup: two per line of a pass where the 8 N^2 of them fit
// a workgroup (N <= 11), each computing the N outputs of one of the line's
// two components, else one per line (N = 12), in whole waves, so that each
// pass is one round and each thread's chain of FMAs is short (one Helmholtz
// apply at k = 9, 2688 output nodes, when every level ran this kernel: post
// transforms 0.62 ms with 256 threads, 0.42 ms with one thread per line,
// 0.36 ms with two).  These kernels take the levels with more branches than
// k_post_one below is used for at N <= 10; at N = 11, 12 k_post_two takes
// them.
template <int N>
struct PostShape {
    static constexpr int H = (8 * N * N <= 1024) ? 2 : 1;  // threads per line
    static constexpr int T = 64 * ((H * 4 * N * N + 63) / 64);
};

template <int N, int DIM>
__device__ __forceinline__ int post_index(int a, int c, int i) {
    if constexpr (DIM == 0) return a * N + c * N * N + i;
    else if constexpr (DIM == 1) return a + c * N * N + i * N;
    else return a + c * N + i * N * N;
}

template <int N, int DIM>
__device__ __forceinline__ void post_line_pass(double *blk, const double *MT, int tid) {
    constexpr int N3 = N * N * N, NN = N * N, N2 = 2 * N, bit = 1 << DIM, low = bit - 1;
    constexpr int H = PostShape<N>::H, NO = N2 / H;  // outputs per thread
    static_assert(PostShape<N>::T >= H * 4 * NN && PostShape<N>::T <= 1024, "one line (half) per thread");
    // lanes 2i and 2i + 1 share line i (their input reads are one address)
    const bool live = tid < H * 4 * NN;
    const int L = live ? tid / H : 0, h = live ? tid % H : 0;
    const int pr = L / NN, ac = L % NN, a = ac % N, c = ac / N;
    const int t0 = ((pr & ~low) << 1) | (pr & low), t1 = t0 | bit;
    double o[NO];
    if (live) {
        // each output a sum over j = 0 .. 2N - 1 in turn
#pragma unroll
        for (int r = 0; r < NO; r++) o[r] = 0.0;
        if constexpr (H > 1) {
            // the line's 2N inputs first (all loads in flight at once)
            double v[N2];
#pragma unroll
            for (int j = 0; j < N2; j++)
                v[j] = blk[(j < N ? t0 : t1) * N3 + post_index<N, DIM>(a, c, j < N ? j : j - N)];
#pragma unroll
            for (int j = 0; j < N2; j++) {
#pragma unroll
                for (int r = 0; r < NO; r++) o[r] = fma(MT[j * N2 + h * NO + r], v[j], o[r]);
            }
        } else {
            // N = 12: one input at a time into the 2N accumulators, which
            // keeps the kernel free of spills
#pragma unroll 1
            for (int j = 0; j < N2; j++) {
                const double vj = blk[(j < N ? t0 : t1) * N3 + post_index<N, DIM>(a, c, j < N ? j : j - N)];
#pragma unroll
                for (int r = 0; r < NO; r++) o[r] = fma(MT[j * N2 + r], vj, o[r]);
            }
        }
    }
    // with two threads per line, every input of the pass is read before
    // any output is written back in its place
    if constexpr (H > 1) __syncthreads();
    if (live) {
        if constexpr (H == 1) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                blk[t0 * N3 + post_index<N, DIM>(a, c, i)] = o[i];
                blk[t1 * N3 + post_index<N, DIM>(a, c, i)] = o[N + i];
            }
        } else {
            const int t = h ? t1 : t0;
#pragma unroll
            for (int i = 0; i < N; i++) blk[t * N3 + post_index<N, DIM>(a, c, i)] = o[i];
        }
    }
}

// Block element e of a branch's 8 N^3 values, for the thread's e in turn:
// e = tid + k T.  Where a thread has at most 10 of them (N <= 10) the loop is
// unrolled whole, so that all of a thread's global loads are in flight at
// once (the loads of one level launch wait on memory only once).
template <int N>
struct PostBlock {
    static constexpr int E = 8 * N * N * N, T = PostShape<N>::T, K = (E + T - 1) / T;
    static constexpr bool whole = K <= 10;
};

// blk[e] = src[map(e)] for the block's elements
template <int N, class Map>
__device__ __forceinline__ void post_gather(double *blk, const double *src, int tid, Map map) {
    using B = PostBlock<N>;
    if constexpr (B::whole) {
        double v[B::K];
#pragma unroll
        for (int k = 0; k < B::K; k++) {
            const int e = tid + k * B::T;
            v[k] = e < B::E ? src[map(e)] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < B::K; k++) {
            const int e = tid + k * B::T;
            if (e < B::E) blk[e] = v[k];
        }
    } else {
#pragma unroll 8
        for (int e = tid; e < B::E; e += B::T) blk[e] = src[map(e)];
    }
}

// TopDown step for one level: each branch's block reconstructed into its 8
// children's scaling parts (directions z, y, x), added to theirs.
template <int N>
__global__ __launch_bounds__(PostShape<N>::T) void k_topdown(const PostPair *pairs, int n_pairs,
                                                          const double *__restrict__ F) {
    constexpr int N3 = N * N * N, N2 = 2 * N;
    __shared__ double blk[8 * N3];
    __shared__ double M[N2 * N2];  // transposed matrix of F^T: F itself
    const int tid = threadIdx.x;
    for (int e = tid; e < N2 * N2; e += PostShape<N>::T) M[e] = F[e];
    for (int w = blockIdx.x; w < n_pairs; w += gridDim.x) {
        const PostPair pr = pairs[w];
        // unrolled, so that several global loads are in flight per thread
        post_gather<N>(blk, pr.node, tid, [](int e) { return (size_t)e; });
        __syncthreads();
        post_line_pass<N, 2>(blk, M, tid);
        __syncthreads();
        post_line_pass<N, 1>(blk, M, tid);
        __syncthreads();
        post_line_pass<N, 0>(blk, M, tid);
        __syncthreads();
        // the children's scaling parts += the reconstruction: every load
        // of the thread first, then the stores
        if constexpr (PostBlock<N>::whole) {
            using B = PostBlock<N>;
            double v[B::K];
#pragma unroll
            for (int k = 0; k < B::K; k++) {
                const int e = tid + k * B::T;
                v[k] = e < B::E ? pr.child0[(size_t)(e / N3) * 8 * N3 + e % N3] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < B::K; k++) {
                const int e = tid + k * B::T;
                if (e < B::E) pr.child0[(size_t)(e / N3) * 8 * N3 + e % N3] = v[k] + blk[e];
            }
        } else {
#pragma unroll 8
            for (int e = tid; e < 8 * N3; e += PostShape<N>::T) pr.child0[(size_t)(e / N3) * 8 * N3 + e % N3] += blk[e];
        }
        __syncthreads();
    }
}

// BottomUp step for one level: each branch's [s, d1..d7] from its children's
// scaling parts (directions x, y, z).
template <int N>
__global__ __launch_bounds__(PostShape<N>::T) void k_compress(const PostPair *pairs, int n_pairs,
                                                           const double *__restrict__ F) {
    constexpr int N3 = N * N * N, N2 = 2 * N;
    __shared__ double blk[8 * N3];
    __shared__ double M[N2 * N2];  // transposed matrix of F: F^T
    const int tid = threadIdx.x;
    for (int e = tid; e < N2 * N2; e += PostShape<N>::T) M[e] = F[(e % N2) * N2 + e / N2];
    for (int w = blockIdx.x; w < n_pairs; w += gridDim.x) {
        const PostPair pr = pairs[w];
        post_gather<N>(blk, pr.child0, tid, [](int e) { return (size_t)(e / N3) * 8 * N3 + e % N3; });
        __syncthreads();
        post_line_pass<N, 0>(blk, M, tid);
        __syncthreads();
        post_line_pass<N, 1>(blk, M, tid);
        __syncthreads();
        post_line_pass<N, 2>(blk, M, tid);
        __syncthreads();
#pragma unroll 8
        for (int e = tid; e < 8 * N3; e += PostShape<N>::T) pr.node[e] = blk[e];
        __syncthreads();
    }
}

// Levels with few branches: one workgroup per (branch, output block), so
// that a branch's three passes spread over 8 CUs.  TopDown: output block ch
// is child ch's reconstructed scaling part (added into it); BottomUp: output
// block t is component t of the branch's [s, d1..d7].  Each pass keeps only
// the half of its outputs that leads to the workgroup's block: the first
// pass reads its lines from global memory and leaves 4 N^3 values in LDS,
// the second 2 N^3, the third N^3, which go straight to global memory.
// Every value kept is the same FMA sequence as in k_topdown / k_compress,
// so the results are the same to the bit.  The 8 workgroups of a branch
// have equal blockIdx modulo 8, so they run on one XCD and share its L2:
// blockIdx = 64 (p / 8) + 8 ch + p % 8 for branch p and block ch.
template <int N>
struct PostOneShape {
    static constexpr int T = 64 * ((4 * N * N + 63) / 64);  // one thread per line of the first pass
    static constexpr int LDS = (4 * N * N * N + 4 * N * N) * 8;
    // workgroups resident per CU by LDS (at most 4), and the largest level
    // that then fits the 256 CUs in one round
    static constexpr int PER_CU = (160 * 1024 / LDS) < 4 ? (160 * 1024 / LDS) : 4;
    static constexpr int MAX_PAIRS = 256 * PER_CU / 8;
};

// o[i] = sum over j of MT[j][h N + i] v[j], j = 0 .. 2N - 1 in turn
template <int N>
__device__ __forceinline__ void post_half_line(const double *v, const double *MT, int h, double *o) {
    constexpr int N2 = 2 * N;
#pragma unroll
    for (int i = 0; i < N; i++) o[i] = 0.0;
#pragma unroll
    for (int j = 0; j < N2; j++) {
#pragma unroll
        for (int i = 0; i < N; i++) o[i] = fma(MT[j * N2 + h * N + i], v[j], o[i]);
    }
}

template <int N, bool DOWN>
__global__ __launch_bounds__(PostOneShape<N>::T) void k_post_one(const PostPair *pairs, int n_pairs,
                                                              const double *__restrict__ F) {
    constexpr int N3 = N * N * N, NN = N * N, N2 = 2 * N;
    __shared__ double X[4 * N3];  // slots of the first pass's outputs
    __shared__ double M[N2 * N2];
    const int tid = threadIdx.x, b = blockIdx.x;
    const int p = (b >> 6) * 8 + (b & 7), ob = (b >> 3) & 7;  // branch, output block
    if (p >= n_pairs) return;
    // TopDown reads F as is, BottomUp its transpose (as k_topdown / k_compress)
    for (int e = tid; e < N2 * N2; e += PostOneShape<N>::T) M[e] = DOWN ? F[e] : F[(e % N2) * N2 + e / N2];
    const PostPair pr = pairs[p];
    const int hx = ob & 1, hy = (ob >> 1) & 1, hz = ob >> 2;
    double v[N2], o[N];
    // pass 1 from global memory: TopDown along z over the branch's 8 blocks
    // (slot s = (ty, tx), blocks s and s | 4), BottomUp along x over the
    // children's scaling parts (slot s = (cz, cy), children 2s and 2s + 1)
    const bool l1 = tid < 4 * NN;
    const int s1 = tid / NN, ac = tid % NN, a = ac % N, c = ac / N;
    if (l1) {
#pragma unroll
        for (int j = 0; j < N2; j++) {
            const int jj = j < N ? j : j - N;
            if constexpr (DOWN) v[j] = pr.node[(j < N ? s1 : s1 | 4) * N3 + post_index<N, 2>(a, c, jj)];
            else v[j] = pr.child0[(size_t)(j < N ? 2 * s1 : 2 * s1 + 1) * 8 * N3 + post_index<N, 0>(a, c, jj)];
        }
    }
    __syncthreads();  // M in LDS
    if (l1) {
        post_half_line<N>(v, M, DOWN ? hz : hx, o);
#pragma unroll
        for (int i = 0; i < N; i++) X[s1 * N3 + (DOWN ? post_index<N, 2>(a, c, i) : post_index<N, 0>(a, c, i))] = o[i];
    }
    __syncthreads();
    // pass 2 along y: TopDown over slots (tx, tx | 2) into tx | 2 hy,
    // BottomUp over slots (2 cz, 2 cz + 1) into 2 cz + hy
    if (tid < 2 * NN) {
        const int q = tid / NN;
        const int s0 = DOWN ? q : 2 * q, s1b = DOWN ? (q | 2) : (2 * q + 1);
#pragma unroll
        for (int j = 0; j < N2; j++) v[j] = X[(j < N ? s0 : s1b) * N3 + post_index<N, 1>(a, c, j < N ? j : j - N)];
        post_half_line<N>(v, M, hy, o);
        const int so = DOWN ? (q | (hy << 1)) : (2 * q + hy);
#pragma unroll
        for (int i = 0; i < N; i++) X[so * N3 + post_index<N, 1>(a, c, i)] = o[i];
    }
    __syncthreads();
    // pass 3: TopDown along x over slots (2 hy, 2 hy + 1), added into child
    // ob's scaling part; BottomUp along z over slots (hy, hy | 2), stored as
    // block ob of the branch's node
    if (tid < NN) {
        const int s0 = DOWN ? 2 * hy : hy, s1b = DOWN ? 2 * hy + 1 : (hy | 2);
        constexpr int D3 = DOWN ? 0 : 2;
#pragma unroll
        for (int j = 0; j < N2; j++) v[j] = X[(j < N ? s0 : s1b) * N3 + post_index<N, D3>(a, c, j < N ? j : j - N)];
        post_half_line<N>(v, M, DOWN ? hx : hz, o);
        if constexpr (DOWN) {
            double *dst = pr.child0 + (size_t)ob * 8 * N3;
            double old[N];
#pragma unroll
            for (int i = 0; i < N; i++) old[i] = dst[post_index<N, 0>(a, c, i)];
#pragma unroll
            for (int i = 0; i < N; i++) dst[post_index<N, 0>(a, c, i)] = old[i] + o[i];
        } else {
#pragma unroll
            for (int i = 0; i < N; i++) pr.node[ob * N3 + post_index<N, 2>(a, c, i)] = o[i];
        }
    }
}

// Wide levels at N >= 11, where k_topdown / k_compress hold 8 N^3 values in
// LDS and so run one workgroup per CU: one workgroup per (branch, half h of
// the first pass's outputs).  The first pass reads its lines from global
// memory and keeps half h (4 N^3 values in LDS: half the LDS, two or three
// workgroups per CU), the second and third passes run whole on those, two
// threads per line, and the workgroup's 4 output blocks go to global memory
// (TopDown: the children with bit h along z, added into their scaling parts;
// BottomUp: the node's components with bit h along x).  No FMA is added, and
// every value is the same FMA sequence as in k_topdown / k_compress.  The two
// workgroups of a branch share an XCD: blockIdx = 16 (p / 8) + 8 h + p % 8.
template <int N>
struct PostTwoShape {
    static constexpr int T = 64 * ((4 * N * N + 63) / 64);
};

// A whole pass along DIM over slot pairs (s0, s0 | sbit), s0 = q << qshift
// for q = 0, 1: two threads per line, each the N outputs of one slot.
template <int N, int DIM>
__device__ __forceinline__ void post_slot_pass(double *X, const double *MT, int tid, int sbit, int qshift) {
    constexpr int N3 = N * N * N, NN = N * N, N2 = 2 * N;
    const bool live = tid < 4 * NN;
    const int L = live ? tid >> 1 : 0, h = tid & 1;
    const int q = L / NN, ac = L % NN, a = ac % N, c = ac / N;
    const int s0 = q << qshift, s1 = s0 | sbit;
    double v[N2], o[N];
    if (live) {
#pragma unroll
        for (int j = 0; j < N2; j++) v[j] = X[(j < N ? s0 : s1) * N3 + post_index<N, DIM>(a, c, j < N ? j : j - N)];
        post_half_line<N>(v, MT, h, o);
    }
    __syncthreads();  // both slots of every line read before any is written
    if (live) {
        const int so = h ? s1 : s0;
#pragma unroll
        for (int i = 0; i < N; i++) X[so * N3 + post_index<N, DIM>(a, c, i)] = o[i];
    }
}

template <int N, bool DOWN>
__global__ __launch_bounds__(PostTwoShape<N>::T) void k_post_two(const PostPair *pairs, int n_pairs,
                                                              const double *__restrict__ F) {
    constexpr int N3 = N * N * N, NN = N * N, N2 = 2 * N;
    __shared__ double X[4 * N3];
    __shared__ double M[N2 * N2];
    const int tid = threadIdx.x, b = blockIdx.x;
    const int p = (b >> 4) * 8 + (b & 7), h0 = (b >> 3) & 1;
    if (p >= n_pairs) return;
    for (int e = tid; e < N2 * N2; e += PostTwoShape<N>::T) M[e] = DOWN ? F[e] : F[(e % N2) * N2 + e / N2];
    const PostPair pr = pairs[p];
    // pass 1 from global memory, as in k_post_one: TopDown along z (slot
    // s = (ty, tx), blocks s and s | 4), BottomUp along x (slot s = (cz, cy),
    // children 2s and 2s + 1); half h0 kept
    double v[N2], o[N];
    const bool l1 = tid < 4 * NN;
    const int s1 = tid / NN, ac = tid % NN, a = ac % N, c = ac / N;
    if (l1) {
#pragma unroll
        for (int j = 0; j < N2; j++) {
            const int jj = j < N ? j : j - N;
            if constexpr (DOWN) v[j] = pr.node[(j < N ? s1 : s1 | 4) * N3 + post_index<N, 2>(a, c, jj)];
            else v[j] = pr.child0[(size_t)(j < N ? 2 * s1 : 2 * s1 + 1) * 8 * N3 + post_index<N, 0>(a, c, jj)];
        }
    }
    __syncthreads();  // M in LDS
    if (l1) {
        post_half_line<N>(v, M, h0, o);
#pragma unroll
        for (int i = 0; i < N; i++) X[s1 * N3 + (DOWN ? post_index<N, 2>(a, c, i) : post_index<N, 0>(a, c, i))] = o[i];
    }
    __syncthreads();
    if constexpr (DOWN) {
        post_slot_pass<N, 1>(X, M, tid, 2, 0);  // y: slots (tx, tx | 2)
        __syncthreads();
        post_slot_pass<N, 0>(X, M, tid, 1, 1);  // x: slots (2 ty, 2 ty + 1)
        __syncthreads();
        // slot s = (ty, tx) is child s | 4 h0: added into its scaling part
        for (int e = tid; e < 4 * N3; e += PostTwoShape<N>::T) {
            double *dst = pr.child0 + (size_t)((e / N3) | (h0 << 2)) * 8 * N3 + e % N3;
            *dst = *dst + X[e];
        }
    } else {
        post_slot_pass<N, 1>(X, M, tid, 1, 1);  // y: slots (2 cz, 2 cz + 1)
        __syncthreads();
        post_slot_pass<N, 2>(X, M, tid, 2, 0);  // z: slots (ty, ty | 2)
        __syncthreads();
        // slot s = (tz, ty) is component 2 s + h0 of the node
        for (int e = tid; e < 4 * N3; e += PostTwoShape<N>::T)
            pr.node[(size_t)(2 * (e / N3) + h0) * N3 + e % N3] = X[e];
    }
}

// One level of the TopDown (down) or BottomUp sequence.
template <int N>
static hipError_t post_level_n(hipStream_t s, bool down, const PostPair *pairs, int cnt, const double *F) {
    if (cnt <= PostOneShape<N>::MAX_PAIRS) {
        const unsigned g = (unsigned)((cnt + 7) / 8) * 64;
        if (down)
            hipLaunchKernelGGL((k_post_one<N, true>), dim3(g), dim3(PostOneShape<N>::T), 0, s, pairs, cnt, F);
        else
            hipLaunchKernelGGL((k_post_one<N, false>), dim3(g), dim3(PostOneShape<N>::T), 0, s, pairs, cnt, F);
        return hipGetLastError();
    }
    if constexpr (N >= 11) {
        const unsigned g = (unsigned)((cnt + 7) / 8) * 16;
        if (down)
            hipLaunchKernelGGL((k_post_two<N, true>), dim3(g), dim3(PostTwoShape<N>::T), 0, s, pairs, cnt, F);
        else
            hipLaunchKernelGGL((k_post_two<N, false>), dim3(g), dim3(PostTwoShape<N>::T), 0, s, pairs, cnt, F);
        return hipGetLastError();
    }
    const int g = std::min(cnt, kMulBlocks);
    if (down)
        hipLaunchKernelGGL((k_topdown<N>), dim3(g), dim3(PostShape<N>::T), 0, s, pairs, cnt, F);
    else
        hipLaunchKernelGGL((k_compress<N>), dim3(g), dim3(PostShape<N>::T), 0, s, pairs, cnt, F);
    return hipGetLastError();
}

static hipError_t post_level(hipStream_t s, int n, bool down, const PostPair *pairs, int cnt, const double *F) {
    switch (n) {
        case 2: return post_level_n<2>(s, down, pairs, cnt, F);
        case 3: return post_level_n<3>(s, down, pairs, cnt, F);
        case 4: return post_level_n<4>(s, down, pairs, cnt, F);
        case 5: return post_level_n<5>(s, down, pairs, cnt, F);
        case 6: return post_level_n<6>(s, down, pairs, cnt, F);
        case 7: return post_level_n<7>(s, down, pairs, cnt, F);
        case 8: return post_level_n<8>(s, down, pairs, cnt, F);
        case 9: return post_level_n<9>(s, down, pairs, cnt, F);
        case 10: return post_level_n<10>(s, down, pairs, cnt, F);
        case 11: return post_level_n<11>(s, down, pairs, cnt, F);
        case 12: return post_level_n<12>(s, down, pairs, cnt, F);
        default: return hipErrorInvalidValue;
    }
}

int dev_pool_fill(DeviceCtx *c, const MRA &mra, const std::vector<PoolFill> &jobs, std::string &err) {
    if (jobs.empty()) return 0;
    HIP_TRY(hipSetDevice(dev_device(c)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(c));
    const int n = mra.kp1(), n3 = n * n * n;
    const size_t nj = jobs.size();
    // generated nodes of all jobs by depth (a job's depth lv after its depth
    // lv - 1), then every job's slots
    size_t nlv = 0;
    std::vector<size_t> gen_at(nj + 1, 0), slot_at(nj + 1, 0);
    for (size_t i = 0; i < nj; i++) {
        const ApplyPlan &pl = *jobs[i].pl;
        nlv = std::max(nlv, pl.gen_level_off.empty() ? (size_t)0 : pl.gen_level_off.size() - 1);
        gen_at[i + 1] = gen_at[i] + pl.gen_parent.size();
        slot_at[i + 1] = slot_at[i] + pl.slot_src.size();
    }
    const size_t n_gen = gen_at[nj], n_slot = slot_at[nj];
    if (n_gen >= (1u << 31) || n_slot >= (1u << 31)) {
        err = "pool fill: too many slots";
        return -1;
    }
    // parent groups (runs of one parent within a depth: the plan orders the
    // children of a parent next to each other) of depth lv:
    // [lv_at[lv], lv_at[lv + 1])
    auto n_groups = [&](const ApplyPlan &pl, size_t lv) {
        size_t k = 0;
        for (int g = pl.gen_level_off[lv]; g < pl.gen_level_off[lv + 1]; g++)
            k += (g == pl.gen_level_off[lv] || pl.gen_parent[g] != pl.gen_parent[g - 1]);
        return k;
    };
    std::vector<size_t> lv_at(nlv + 1, 0);
    for (size_t lv = 0; lv < nlv; lv++) {
        size_t k = 0;
        for (size_t i = 0; i < nj; i++)
            if (lv + 1 < jobs[i].pl->gen_level_off.size()) k += n_groups(*jobs[i].pl, lv);
        lv_at[lv + 1] = lv_at[lv] + k;
    }
    const size_t n_grp = lv_at[nlv];
    // one device allocation: [filter | gen items | slot items] (uploaded)
    // [generated scaling blocks of every job] [every slot's norms]
    size_t bytes = 0;
    auto place = [&](size_t b) {
        const size_t off = bytes;
        bytes += (std::max<size_t>(b, 8) + 255) / 256 * 256;
        return off;
    };
    const size_t f_F = place(sizeof(double) * mra.basis.F.size());
    const size_t f_gi = place(sizeof(GenGroup) * n_grp), f_si = place(sizeof(SlotItem) * n_slot);
    const size_t up_bytes = bytes;
    const size_t f_gen = place(sizeof(double) * n_gen * (size_t)n3), f_nrm = place(sizeof(double) * 8 * n_slot);
    unsigned char *dev = nullptr;
    HIP_TRY(hipMalloc((void **)&dev, bytes));
    double *gen_s = reinterpret_cast<double *>(dev + f_gen), *norms = reinterpret_cast<double *>(dev + f_nrm);
    std::vector<unsigned char, UninitAlloc<unsigned char>> stage(up_bytes);
    std::memcpy(stage.data() + f_F, mra.basis.F.data(), sizeof(double) * mra.basis.F.size());
    GenGroup *gi = reinterpret_cast<GenGroup *>(stage.data() + f_gi);
    SlotItem *si = reinterpret_cast<SlotItem *>(stage.data() + f_si);
    // job i's generated node g is gen_s block gen_at[i] + g
    std::vector<size_t> lv_fill(lv_at.begin(), lv_at.end() - (nlv ? 1 : 0));
    for (size_t i = 0; i < nj; i++) {
        const ApplyPlan &pl = *jobs[i].pl;
        const double *inp = jobs[i].inp;
        double *gs = gen_s + gen_at[i] * (size_t)n3;
        for (size_t lv = 0; lv + 1 < pl.gen_level_off.size(); lv++) {
            GenGroup *cur = nullptr;
            for (int g = pl.gen_level_off[lv]; g < pl.gen_level_off[lv + 1]; g++) {
                const int par = pl.gen_parent[g];
                if (g == pl.gen_level_off[lv] || par != pl.gen_parent[g - 1]) {
                    cur = &gi[lv_fill[lv]++];
                    cur->src = par < 0 ? inp + (size_t)(-par - 1) * 8 * n3 : gs + (size_t)par * n3;
                    for (int t = 0; t < 8; t++) cur->out[t] = nullptr;
                    cur->full = par < 0;
                    cur->mask = 0;
                }
                const int t = pl.gen_child[g];
                if ((cur->mask >> t) & 1) {
                    (void)hipFree(dev);
                    err = "pool fill: a generated node twice";
                    return -1;
                }
                cur->out[t] = gs + (size_t)g * n3;
                cur->mask |= 1 << t;
            }
        }
    }
    for (size_t i = 0; i < nj; i++) {
        const ApplyPlan &pl = *jobs[i].pl;
        const double *inp = jobs[i].inp;
        const double *gs = gen_s + gen_at[i] * (size_t)n3;
        for (size_t q = 0; q < pl.slot_src.size(); q++) {
            const int sr = pl.slot_src[q];
            SlotItem &it = si[slot_at[i] + q];
            it.src = sr >= 0 ? inp + (size_t)sr * 8 * n3 : gs + (size_t)(-sr - 1) * n3;
            it.dst = jobs[i].pool + pl.slot_off[q];
            it.norms = norms + (slot_at[i] + q) * 8;
            it.len = sr >= 0 ? 8 * n3 : n3;
        }
    }
    std::vector<double, UninitAlloc<double>> norms_h(8 * n_slot);
    const double *F = reinterpret_cast<const double *>(dev + f_F);
    const GenGroup *dgi = reinterpret_cast<const GenGroup *>(dev + f_gi);
    const SlotItem *dsi = reinterpret_cast<const SlotItem *>(dev + f_si);
    hipError_t e = hipMemcpyAsync(dev, stage.data(), up_bytes, hipMemcpyHostToDevice, s);
    for (size_t lv = 0; e == hipSuccess && lv < nlv; lv++) {
        const int cnt = (int)(lv_at[lv + 1] - lv_at[lv]);
        if (cnt == 0) continue;
        hipLaunchKernelGGL(k_gen_fill, dim3(std::min(cnt, kMulBlocks)), dim3(kGenThreads), sizeof(double) * 6 * n3, s,
                           n, F, dgi + lv_at[lv], cnt);
        e = hipGetLastError();
    }
    if (e == hipSuccess && n_slot > 0) {
        hipLaunchKernelGGL(k_pool_fill, dim3((unsigned)std::min<size_t>(n_slot, 16384)), dim3(64), 0, s, dsi,
                           (int)n_slot);
        e = hipGetLastError();
    }
    if (e == hipSuccess && n_slot > 0) {
        hipLaunchKernelGGL(k_slot_norms, dim3((unsigned)((8 * n_slot + 255) / 256)), dim3(256), 0, s, n, dsi,
                           (int)n_slot);
        e = hipGetLastError();
    }
    if (e == hipSuccess && n_slot > 0)
        e = hipMemcpyAsync(norms_h.data(), norms, sizeof(double) * 8 * n_slot, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dev);
    if (e != hipSuccess) {
        err = std::string("pool fill: ") + hipGetErrorString(e);
        return -2;
    }
    for (size_t i = 0; i < nj; i++) {
        ApplyPlan &pl = *jobs[i].pl;
        const size_t cnt = pl.slot_src.size() * 8;
        if (cnt) std::memcpy(pl.in_norms.data(), norms_h.data() + slot_at[i] * 8, sizeof(double) * cnt);
    }
    return 0;
}

// Block src_idx[i] of src (blk doubles; i itself without src_idx) to block
// dst_idx[i] of dst (i without dst_idx): one workgroup per block.
__global__ __launch_bounds__(256) void k_move_blocks(const double *src, const int *src_idx, const int *dst_idx, int n,
                                                     int blk, double *dst) {
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const double *a = src + (size_t)(src_idx ? src_idx[i] : i) * blk;
        double *b = dst + (size_t)(dst_idx ? dst_idx[i] : i) * blk;
        for (int e = threadIdx.x; e < blk; e += 256) b[e] = a[e];
    }
}

__global__ __launch_bounds__(256) void k_scale(double *p, long long n, double f) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] *= f;
}

int dev_scale(DeviceCtx *c, double *p, size_t n, double f, std::string &err) {
    if (n == 0) return 0;
    HIP_TRY(hipSetDevice(dev_device(c)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(c));
    hipLaunchKernelGGL(k_scale, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0, s, p,
                       (long long)n, f);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int dev_alloc(DeviceCtx *c, size_t n, double **dst, std::string &err) {
    HIP_TRY(hipSetDevice(dev_device(c)));
    *dst = nullptr;
    HIP_TRY(hipMalloc((void **)dst, sizeof(double) * std::max<size_t>(n, 1)));
    return 0;
}

static int move_blocks(DeviceCtx *c, const double *src, const std::vector<int> *src_idx,
                       const std::vector<int> *dst_idx, size_t n, int blk, double *dst, std::string &err) {
    if (n == 0) return 0;
    HIP_TRY(hipSetDevice(dev_device(c)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(c));
    int *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(int) * n));
    const std::vector<int> &idx = src_idx ? *src_idx : *dst_idx;
    hipError_t e = hipMemcpyAsync(d, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_move_blocks, dim3((unsigned)std::min<size_t>(n, 4096)), dim3(256), 0, s, src,
                           src_idx ? d : nullptr, src_idx ? nullptr : d, (int)n, blk, dst);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) {
        err = std::string("block move: ") + hipGetErrorString(e);
        return -2;
    }
    return 0;
}

int dev_scatter_blocks(DeviceCtx *c, const double *src, const std::vector<int> &dst_idx, int blk, double *dst,
                       std::string &err) {
    return move_blocks(c, src, nullptr, &dst_idx, dst_idx.size(), blk, dst, err);
}

int dev_gather_blocks(DeviceCtx *c, const double *src, const std::vector<int> &src_idx, int blk, double *dst,
                      std::string &err) {
    return move_blocks(c, src, &src_idx, nullptr, src_idx.size(), blk, dst, err);
}

int dev_upload_into(DeviceCtx *c, const double *src, size_t n, double *dst, std::string &err) {
    return dev_copy_h2d(c, dst, src, sizeof(double) * n, err);
}

int dev_d2d(DeviceCtx *c, double *dst, const double *src, size_t n, std::string &err) {
    if (n == 0) return 0;
    HIP_TRY(hipSetDevice(dev_device(c)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(c));
    HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int dev_upload(DeviceCtx *c, const double *src, size_t n, double **dst, std::string &err) {
    HIP_TRY(hipSetDevice(dev_device(c)));
    *dst = nullptr;
    HIP_TRY(hipMalloc((void **)dst, sizeof(double) * std::max<size_t>(n, 1)));
    const int rc = dev_copy_h2d(c, *dst, src, sizeof(double) * n, err);
    if (rc) {
        (void)hipFree(*dst);
        *dst = nullptr;
        err = "upload: " + err;
    }
    return rc;
}

int dev_node_norms(DeviceCtx *c, int kp1, const double *blocks, long long n, double *host_s2, double *host_w2,
                   std::string &err) {
    if (n <= 0) return 0;
    HIP_TRY(hipSetDevice(dev_device(c)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(c));
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(double) * 2 * (size_t)n));
    hipLaunchKernelGGL(k_node_norms, dim3((unsigned)n), dim3(256), 0, s, blocks, n, kp1 * kp1 * kp1, d, d + n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host_s2, d, sizeof(double) * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(host_w2, d + n, sizeof(double) * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) {
        err = std::string("node norms: ") + hipGetErrorString(e);
        return -2;
    }
    return 0;
}

int dev_component_norms(DeviceCtx *c, int kp1, const double *blocks, long long n, double *host8, std::string &err) {
    if (n <= 0) return 0;
    HIP_TRY(hipSetDevice(dev_device(c)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(c));
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(double) * 8 * (size_t)n));
    hipLaunchKernelGGL(k_component_norms, dim3((unsigned)n), dim3(64), 0, s, blocks, n, kp1 * kp1 * kp1, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host8, d, sizeof(double) * 8 * (size_t)n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) {
        err = std::string("component norms: ") + hipGetErrorString(e);
        return -2;
    }
    return 0;
}

int dev_post_transform(DeviceCtx *c, const MRA &mra, const std::vector<std::vector<PostPair>> &levels,
                       std::string &err, bool topdown) {
    size_t most = 0, total = 0;
    for (const auto &l : levels) {
        most = std::max(most, l.size());
        total += l.size();
    }
    if (most == 0) return 0;
    if (mra.kp1() > 12) {
        err = "post transforms: unsupported basis size";
        return -1;
    }
    HIP_TRY(hipSetDevice(dev_device(c)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(c));
    double *F = nullptr;
    PostPair *pairs = nullptr;
    std::vector<PostPair> all;
    std::vector<long long> off(1, 0);
    all.reserve(total);
    for (const auto &l : levels) {
        all.insert(all.end(), l.begin(), l.end());
        off.push_back((long long)all.size());
    }
    hipError_t e = hipMalloc((void **)&F, sizeof(double) * mra.basis.F.size());
    if (e == hipSuccess) e = hipMalloc((void **)&pairs, sizeof(PostPair) * all.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(F, mra.basis.F.data(), sizeof(double) * mra.basis.F.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(pairs, all.data(), sizeof(PostPair) * all.size(), hipMemcpyHostToDevice, s);
    int rc = 0;
    if (e == hipSuccess) rc = dev_post_launch(s, mra.kp1(), pairs, off, F, err, topdown);
    if (e == hipSuccess && rc == 0) e = hipStreamSynchronize(s);
    for (void *p : {(void *)F, (void *)pairs})
        if (p) (void)hipFree(p);
    if (rc) return rc;
    if (e != hipSuccess) {
        err = std::string("post transforms: ") + hipGetErrorString(e);
        return -2;
    }
    return 0;
}

int dev_post_launch(void *stream, int kp1, const PostPair *pairs, const std::vector<long long> &lvl_off,
                    const double *F, std::string &err, bool topdown) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nl = (int)lvl_off.size() - 1;
    // TopDown shallow to deep, then BottomUp deep to shallow
    for (int step = topdown ? 0 : nl; step < 2 * nl; step++) {
        const bool down = step < nl;
        const int d = down ? step : 2 * nl - 1 - step;
        const long long cnt = lvl_off[d + 1] - lvl_off[d];
        if (cnt == 0) continue;
        HIP_TRY(post_level(s, kp1, down, pairs + lvl_off[d], (int)cnt, F));
    }
    return 0;
}

int dev_node_norms_launch(void *stream, int kp1, const double *blocks, long long n, double *s2, double *w2,
                          std::string &err) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_node_norms, dim3((unsigned)n), dim3(256), 0, static_cast<hipStream_t>(stream), blocks, n,
                       kp1 * kp1 * kp1, s2, w2);
    HIP_TRY(hipGetLastError());
    return 0;
}

int dev_download(DeviceCtx *c, const double *src, size_t n, double *dst, std::string &err) {
    return dev_copy_d2h(c, dst, src, sizeof(double) * n, err);
}

void dev_free(DeviceCtx *c, void *p) {
    if (!p) return;
    (void)hipSetDevice(dev_device(c));
    (void)hipStreamSynchronize(static_cast<hipStream_t>(dev_stream(c)));
    (void)hipFree(p);
}

// Device memory of the allocation-free product (mwg_multiply_noalloc): one
// array on every GPU the library runs on, held until the process ends.  In
// host code its name is a host-side stand-in: kernels and copies get its
// device address, taken once per device by dev_noalloc_init.
constexpr size_t kNoAllocBytes = (size_t)128 << 20;
constexpr int kNoAllocBlocks = 64;  // workgroups of an allocation-free launch
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

int dev_noalloc_init(DeviceCtx *c, std::string &err) {
    // the device is current (dev_create sets it)
    void *ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    dev_set_noalloc(c, ptr, kNoAllocBytes);
    return 0;
}

struct DevMulSession {
    DeviceCtx *ctx = nullptr;
    int n = 0;
    double *pool_a = nullptr, *pool_b = nullptr, *F = nullptr, *wsc = nullptr, *ws = nullptr;
    // allocation-free: every buffer is carved from the device's scratch
    // array (nothing is freed); output nodes go in chunks of `chunk` nodes
    // through dn / dout
    bool noalloc = false;
    size_t chunk = 0;
    MulNode *dn = nullptr;
    double *dout = nullptr;
    // the pools are the caller's device blocks (not freed here)
    bool borrowed_a = false, borrowed_b = false;
    ~DevMulSession() {
        if (noalloc) return;
        for (double *p : {F, wsc, ws})
            if (p) (void)hipFree(p);
        if (pool_a && !borrowed_a) (void)hipFree(pool_a);
        if (pool_b && !borrowed_b) (void)hipFree(pool_b);
    }
};

int dev_mul_begin(DeviceCtx *c, const MRA &mra, const double *pool_a, size_t n_a, const double *pool_b, size_t n_b,
                  DevMulSession **out, std::string &err, bool noalloc, bool a_on_device, bool b_on_device) {
    HIP_TRY(hipSetDevice(dev_device(c)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(c));
    if (mra.basis.type != BASIS_INTERPOLATING) {
        err = "multiplication needs the interpolating basis";
        return -1;
    }
    const int n = mra.kp1(), n3 = n * n * n;
    std::vector<double> x(n), w(n), wsc(n3);
    const double *qx = nullptr, *qw = nullptr;
    quad_rule01(n, &qx, &qw);
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) wsc[(k * n + j) * n + i] = 1.0 / std::sqrt(qw[i] * qw[j] * qw[k]);
    if (noalloc) {
        // layout in the scratch array (256-byte aligned pieces): filter,
        // point weights, both pools, the workgroups' work areas, then node
        // descriptors and output blocks of one chunk
        unsigned char *base = static_cast<unsigned char *>(dev_noalloc_base(c));
        const size_t cap = dev_noalloc_bytes(c);
        if (!base) {
            err = "no device scratch array on this context";
            return -1;
        }
        auto al = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t fixed = al(sizeof(double) * mra.basis.F.size()) + al(sizeof(double) * n3) +
                             al(sizeof(double) * std::max<size_t>(n_a, 1)) + al(sizeof(double) * std::max<size_t>(n_b, 1)) +
                             al(sizeof(double) * (size_t)kNoAllocBlocks * 24 * n3);
        const size_t per_node = sizeof(MulNode) + sizeof(double) * 8 * (size_t)n3;
        if (fixed + al(per_node) + 256 > cap) {
            err = "mwg_multiply_noalloc: the inputs need " + std::to_string(fixed + per_node) +
                  " bytes of device memory, more than its " + std::to_string(cap) + "-byte scratch array";
            return -1;
        }
        auto *m = new DevMulSession;
        m->ctx = c;
        m->n = n;
        m->noalloc = true;
        size_t off = 0;
        auto take = [&](size_t bytes) {
            unsigned char *p = base + off;
            off += al(bytes);
            return p;
        };
        m->F = reinterpret_cast<double *>(take(sizeof(double) * mra.basis.F.size()));
        m->wsc = reinterpret_cast<double *>(take(sizeof(double) * n3));
        m->pool_a = reinterpret_cast<double *>(take(sizeof(double) * std::max<size_t>(n_a, 1)));
        m->pool_b = reinterpret_cast<double *>(take(sizeof(double) * std::max<size_t>(n_b, 1)));
        m->ws = reinterpret_cast<double *>(take(sizeof(double) * (size_t)kNoAllocBlocks * 24 * n3));
        m->chunk = (cap - off - 256) / per_node;
        m->dn = reinterpret_cast<MulNode *>(take(sizeof(MulNode) * m->chunk));
        m->dout = reinterpret_cast<double *>(base + off);
        hipError_t e = hipMemcpyAsync(m->F, mra.basis.F.data(), sizeof(double) * mra.basis.F.size(),
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(m->wsc, wsc.data(), sizeof(double) * n3, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && n_a)
            e = hipMemcpyAsync(m->pool_a, pool_a, sizeof(double) * n_a,
                               a_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s);
        if (e == hipSuccess && n_b)
            e = hipMemcpyAsync(m->pool_b, pool_b, sizeof(double) * n_b,
                               b_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(s);
            delete m;
            err = std::string("upload: ") + hipGetErrorString(e);
            return -2;
        }
        *out = m;
        return 0;
    }
    auto *m = new DevMulSession;
    m->ctx = c;
    m->n = n;
    auto up = [&](const double *v, size_t nv, double **d) -> int {
        HIP_TRY(hipMalloc((void **)d, sizeof(double) * std::max<size_t>(nv, 1)));
        if (nv) HIP_TRY(hipMemcpyAsync(*d, v, sizeof(double) * nv, hipMemcpyHostToDevice, s));
        return 0;
    };
    int rc = 0;
    if (a_on_device) {
        m->pool_a = const_cast<double *>(pool_a);
        m->borrowed_a = true;
    } else {
        rc = up(pool_a, n_a, &m->pool_a);
    }
    if (!rc && b_on_device) {
        m->pool_b = const_cast<double *>(pool_b);
        m->borrowed_b = true;
    } else if (!rc) {
        rc = up(pool_b, n_b, &m->pool_b);
    }
    if (!rc) rc = up(mra.basis.F.data(), mra.basis.F.size(), &m->F);
    if (!rc) rc = up(wsc.data(), wsc.size(), &m->wsc);
    if (!rc) {
        hipError_t e = hipMalloc((void **)&m->ws, sizeof(double) * (size_t)kMulBlocks * 24 * n3);
        if (e != hipSuccess) {
            err = std::string("hipMalloc(work area): ") + hipGetErrorString(e);
            rc = -2;
        }
    }
    if (rc) {
        (void)hipStreamSynchronize(s);
        delete m;
        return rc;
    }
    *out = m;
    return 0;
}

int dev_mul_nodes(DevMulSession *m, const std::vector<MulNode> &nodes, double coef, double *out_blocks,
                  std::string &err, bool out_on_device) {
    if (nodes.empty()) return 0;
    HIP_TRY(hipSetDevice(dev_device(m->ctx)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(m->ctx));
    const int n3 = m->n * m->n * m->n;
    if (m->noalloc) {
        // chunk by chunk through the scratch array; each output node is one
        // workgroup's, so the blocks equal those of the allocating path
        for (size_t c0 = 0; c0 < nodes.size(); c0 += m->chunk) {
            const size_t cn = std::min(m->chunk, nodes.size() - c0);
            HIP_TRY(hipMemcpyAsync(m->dn, nodes.data() + c0, sizeof(MulNode) * cn, hipMemcpyHostToDevice, s));
            const int grid = (int)std::min<size_t>(cn, kNoAllocBlocks);
            hipLaunchKernelGGL(k_mul_nodes, dim3(grid), dim3(kMulThreads), 0, s, m->n, m->dn, (int)cn, m->pool_a,
                               m->pool_b, m->F, m->wsc, coef, m->ws, m->dout);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(out_blocks + c0 * 8 * n3, m->dout, sizeof(double) * cn * 8 * n3,
                                   out_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        return 0;
    }
    MulNode *dn = nullptr;
    double *dout = out_on_device ? out_blocks : nullptr;
    HIP_TRY(hipMalloc((void **)&dn, sizeof(MulNode) * nodes.size()));
    hipError_t e = out_on_device ? hipSuccess : hipMalloc((void **)&dout, sizeof(double) * nodes.size() * 8 * n3);
    if (e != hipSuccess) {
        (void)hipFree(dn);
        err = std::string("hipMalloc(out): ") + hipGetErrorString(e);
        return -2;
    }
    e = hipMemcpyAsync(dn, nodes.data(), sizeof(MulNode) * nodes.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const int grid = (int)std::min<size_t>(nodes.size(), kMulBlocks);
        hipLaunchKernelGGL(k_mul_nodes, dim3(grid), dim3(kMulThreads), 0, s, m->n, dn, (int)nodes.size(), m->pool_a,
                           m->pool_b, m->F, m->wsc, coef, m->ws, dout);
        e = hipGetLastError();
    }
    int crc = 0;
    if (e == hipSuccess && !out_on_device)
        crc = dev_copy_d2h(m->ctx, out_blocks, dout, sizeof(double) * nodes.size() * 8 * n3, err);
    if (e == hipSuccess && out_on_device) e = hipStreamSynchronize(s);
    (void)hipFree(dn);
    if (!out_on_device) (void)hipFree(dout);
    if (crc) return crc;
    if (e != hipSuccess) {
        err = std::string("multiply launch: ") + hipGetErrorString(e);
        return -2;
    }
    return 0;
}

void dev_mul_end(DevMulSession *m) {
    if (!m) return;
    (void)hipSetDevice(dev_device(m->ctx));
    (void)hipStreamSynchronize(static_cast<hipStream_t>(dev_stream(m->ctx)));
    delete m;
}

struct DevAddSession {
    DeviceCtx *ctx = nullptr;
    int n = 0, n_in = 0;
    double *pool = nullptr, *coefs = nullptr, *F = nullptr, *ws = nullptr;
    ~DevAddSession() {
        for (double *p : {pool, coefs, F, ws})
            if (p) (void)hipFree(p);
    }
};

int dev_add_begin(DeviceCtx *c, const MRA &mra, const std::vector<BlockSource> &pool, const std::vector<double> &coefs,
                  DevAddSession **out, std::string &err) {
    HIP_TRY(hipSetDevice(dev_device(c)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(c));
    auto *m = new DevAddSession;
    m->ctx = c;
    m->n = mra.kp1();
    m->n_in = (int)coefs.size();
    const int n3 = m->n * m->n * m->n;
    auto up = [&](const std::vector<double> &v, double **d) -> int {
        HIP_TRY(hipMalloc((void **)d, sizeof(double) * std::max<size_t>(v.size(), 1)));
        if (!v.empty()) HIP_TRY(hipMemcpyAsync(*d, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, s));
        return 0;
    };
    // the pool: the sources one after another, copied from host or device
    size_t total = 0;
    for (const BlockSource &b : pool) total += b.n;
    int rc = 0;
    {
        hipError_t e = hipMalloc((void **)&m->pool, sizeof(double) * std::max<size_t>(total, 1));
        size_t off = 0;
        for (const BlockSource &b : pool) {
            if (e != hipSuccess) break;
            if (b.n)
                e = hipMemcpyAsync(m->pool + off, b.p, sizeof(double) * b.n,
                                   b.on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s);
            off += b.n;
        }
        if (e != hipSuccess) {
            err = std::string("add pool: ") + hipGetErrorString(e);
            rc = -2;
        }
    }
    if (!rc) rc = up(coefs, &m->coefs);
    if (!rc) rc = up(mra.basis.F, &m->F);
    if (!rc) {
        hipError_t e = hipMalloc((void **)&m->ws, sizeof(double) * (size_t)kMulBlocks * 16 * n3);
        if (e != hipSuccess) {
            err = std::string("hipMalloc(work area): ") + hipGetErrorString(e);
            rc = -2;
        }
    }
    if (rc) {
        (void)hipStreamSynchronize(s);
        delete m;
        return rc;
    }
    *out = m;
    return 0;
}

int dev_add_nodes(DevAddSession *m, const std::vector<AddSrc> &src, double *out_blocks, std::string &err,
                  bool out_on_device) {
    const size_t n_nodes = m->n_in > 0 ? src.size() / (size_t)m->n_in : 0;
    if (n_nodes == 0) return 0;
    if (src.size() != n_nodes * (size_t)m->n_in) {
        err = "source table size is not a multiple of the input count";
        return -1;
    }
    HIP_TRY(hipSetDevice(dev_device(m->ctx)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(m->ctx));
    const int n3 = m->n * m->n * m->n;
    AddSrc *dsrc = nullptr;
    double *dout = out_on_device ? out_blocks : nullptr;
    HIP_TRY(hipMalloc((void **)&dsrc, sizeof(AddSrc) * src.size()));
    hipError_t e = out_on_device ? hipSuccess : hipMalloc((void **)&dout, sizeof(double) * n_nodes * 8 * n3);
    if (e != hipSuccess) {
        (void)hipFree(dsrc);
        err = std::string("hipMalloc(out): ") + hipGetErrorString(e);
        return -2;
    }
    e = hipMemcpyAsync(dsrc, src.data(), sizeof(AddSrc) * src.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const int grid = (int)std::min<size_t>(n_nodes, kMulBlocks);
        hipLaunchKernelGGL(k_add_nodes, dim3(grid), dim3(kMulThreads), 0, s, m->n, dsrc, m->n_in, (int)n_nodes,
                           m->pool, m->coefs, m->F, m->ws, dout);
        e = hipGetLastError();
    }
    int crc = 0;
    if (e == hipSuccess && !out_on_device)
        crc = dev_copy_d2h(m->ctx, out_blocks, dout, sizeof(double) * n_nodes * 8 * n3, err);
    if (e == hipSuccess && out_on_device) e = hipStreamSynchronize(s);
    (void)hipFree(dsrc);
    if (!out_on_device) (void)hipFree(dout);
    if (crc) return crc;
    if (e != hipSuccess) {
        err = std::string("add launch: ") + hipGetErrorString(e);
        return -2;
    }
    return 0;
}

void dev_add_end(DevAddSession *m) {
    if (!m) return;
    (void)hipSetDevice(dev_device(m->ctx));
    (void)hipStreamSynchronize(static_cast<hipStream_t>(dev_stream(m->ctx)));
    delete m;
}

}  // namespace mwg
