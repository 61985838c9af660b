// HIP side of the multiwavelet Green's-function apply (gfx950).
//
// The per-node calculator of upstream MRCPP's OperApplicationCalculator
// (called from mrcpp::apply, src/qmoperators/two_electron/CoulombPotential.cpp:131)
// runs here.  Screening uses the FP64 bound ||O_x|| ||O_y|| ||O_z|| ||f_ft||
// > thr_g evaluated in the same operation order as the host oracle, so the
// set of surviving term-applications is identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
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

constexpr int kMaxActive = 4096;

// Per apply of a batched plan: its operator tables (device copies resident in
// the context) and where its screening tables start.
struct AppDesc {
    const int *opL;                   // [m * ndepth + d]
    const long long *op_boff, *op_noff;
    const double *op_norms;
    const double *blk64;              // FP64 blocks
    const float *blk32;               // FP32 copy (mixed mode), or null
    const _Float16 *blk_h;            // FP16 split rows (split mode), or null
    const double *blk_hsc;            // their per-block scales
    int M, ndepth;
    long long bw_off;                 // start of this apply in bw
    int act_base;                     // start of this apply in act_off
    int mt_d0;                        // first output depth of mtab
    double thr_floor;                 // positive floor of the class threshold
    // pre-screen: per output depth d >= mt_d0, component combination
    // cmb = cx | cy << 2 | cz << 4 and translation |l| <= mt_w, the largest
    // ((N_x N_y) N_z) over the depth's active terms, [d - mt_d0][cmb][lz][ly][lx]
    const double *mtab;
    int mt_w, pad_;
};

struct KArgs {
    // output nodes (all applies, concatenated)
    const int *out_depth;
    const int *out_app;               // apply of each output node
    const double *out_thr;
    const long long *band_off;
    const int *band_w;
    const int *band;
    double *out_ns;
    // input slots (all applies, concatenated)
    const long long *slot_off;
    const double *slot_norm;
    // screening tables
    const int *bw;
    const int *act_off, *act_terms;
    const AppDesc *apps;
    // [0] term-applications passing screening (algorithmic work)
    unsigned long long *counters;
    // work segments: job = 8*out + gt, (band position p, input component ft)
    // pairs 8 p + ft in [p0, p1)
    const int *seg_job, *seg_p0, *seg_p1;
};

template <typename T>
__device__ __forceinline__ const T *op_blocks(const AppDesc &ad);
template <>
__device__ __forceinline__ const double *op_blocks<double>(const AppDesc &ad) {
    return ad.blk64;
}
template <>
__device__ __forceinline__ const float *op_blocks<float>(const AppDesc &ad) {
    return ad.blk32;
}

// One workgroup per job (output node g, output component gt, or the pair of
// components gt and gt | 4 of a pair job); its NW waves share the job's
// 16-wide tiles (TW tiles each).
//
// Screening: the active terms of the job's depth are taken in chunks of 64
// (one term per lane).  For the chunk, the 1D norms of every term, component
// and band translation are staged in LDS.  The (band node f, input component
// ft) pairs are pre-screened 64 at a time against the bound of the depth's
// largest term; for each candidate pair the bound
// ||O_x|| ||O_y|| ||O_z|| ||f_ft|| of every lane-term is formed and a ballot
// gives the surviving terms, which are appended to an LDS worklist of
// (band node, ft, term) items.  A term is taken by this launch when it
// passes the reference rule (bound > thr) and falls in the launch's bound
// class (bound > lo*thr and not bound > hi*thr), so launches with disjoint
// classes together apply exactly the oracle's term set.
//
// Application: each item is three chains of 16x16x4 MFMAs in the launch's
// arithmetic T (FP64 or FP32):
//   x: X[(jz,jy)][ix]  = F[(jz,jy)][jx] . Ox^T[jx][ix]   F tiles in registers
//   y: Y[iy][(jz,ix)]  = Oy[iy][jy] . X'[jy][(jz,ix)]
//   z: acc[iz][(iy,ix)] += Oz[iz][jz] . Y''[jz][(iy,ix)]  acc tiles persist
// A pair job's components share Ox and Oy, so its items run the x and y
// chains once and the z chain for each component whose bound passes (two
// accumulators): 4 chains instead of 6 when both pass.
// X and Y change layout through one padded LDS buffer (two at N >= 11)
// whose lane addresses are a base plus a compile-time offset (padding
// entries stay zero or finite and only meet zero operands).  While an item computes, the operator blocks
// of the next item are loaded (registers, then a double-buffered LDS stage)
// and, when the next item starts a new (f, ft) group, its F tiles too (into
// the F registers, which the x chain no longer needs, during the y and z
// chains).
// MFMA lane layouts (checked on gfx950): A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k; FP64 D[i][j] in lane j + 16 (i mod 4), register i / 4; FP32
// D[i][j] in lane j + 16 (i / 4), register i mod 4.
typedef double mwg_double4 __attribute__((ext_vector_type(4)));
typedef float mwg_float4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef mwg_double4 acc_t;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lq, int r) { return lq + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef mwg_float4 acc_t;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int lq, int r) { return 4 * lq + r; }
};

template <int N>
struct JobShape {
    static constexpr int N2 = N * N, N3 = N2 * N;
    static constexpr int NT = (N2 + 15) / 16;  // 16-wide tiles over N^2
    static constexpr int KS = (N + 3) / 4;     // 4-deep k steps over N
    // waves per job: at most 4 tiles each at N >= 9 (fewer tiles per wave
    // measured slower at N = 10, 12: each wave reads the item's operator
    // operands from LDS once), 2 of 2 tiles at N = 7, 8 (one wave per job
    // left LDS-bound occupancy at 5 waves per CU: 136 exchange applies at
    // k = 7, plain FP64 105.9 -> 97.6 ms)
    static constexpr int NW = N == 7 || N == 8 ? (NT + 1) / 2 : (NT + 3) / 4;
    static constexpr int TW = (NT + NW - 1) / NW;  // tiles per wave
    static constexpr int T = 64 * NW;          // threads per job
    static constexpr int NTP = NW * TW;        // tiles including padding tiles (N = 10: 8 for 7)
    // Xp[row][16], row = (jz,jy) flattened; Yp[16][CP], [iy][(jz,ix)]
    static constexpr int XROWS_W = 16 * NTP;
    static constexpr int XROWS_R = ((16 * NTP - 1) / N) * N + 4 * KS;
    static constexpr int XROWS = XROWS_W > XROWS_R ? XROWS_W : XROWS_R;
    static constexpr int CP = 16 * NTP > 4 * KS * N ? 16 * NTP : 4 * KS * N;
    static constexpr int XSZ = XROWS * 16, YSZ = 16 * CP;
    // N >= 11: X and Y in buffers of their own, so that an item needs two
    // barriers instead of four (64 Helmholtz applies at k = 11, plain FP64:
    // 703 -> 679 ms); at N <= 10 the larger LDS footprint costs a workgroup
    // per CU and one shared buffer is faster
    static constexpr bool SEP = N >= 11;
    static constexpr int BUF = SEP ? XSZ + YSZ : (XSZ > YSZ ? XSZ : YSZ);
    static constexpr int OB = 4 * N2;          // operator values per stage (Ox, Oy, Oz and Oz of gt | 4)
};

constexpr int kMaxBand = 9;     // band half-width limit
constexpr int kWorkCap = 1024;  // worklist items per batch (>= 64 + 512)

// Output job of a work segment and its screening context.
struct JobCtx {
    int g, gt, d, w, side, nl, a0, nact, ndepth;
    long long boff;
    double thr, thr_lo, thr_hi;
    int gb[3];
};

__device__ __forceinline__ JobCtx job_ctx(const KArgs &a, const AppDesc &ad, int job, double lo_mult,
                                          double hi_mult) {
    JobCtx j;
    j.g = job >> 3;
    j.gt = job & 7;
    j.d = a.out_depth[j.g];
    j.thr = a.out_thr[j.g];
    // class cuts relative to the threshold (floored, so that nodes computed
    // without screening still split by magnitude)
    const double thr_cls = j.thr > ad.thr_floor ? j.thr : ad.thr_floor;
    j.thr_lo = thr_cls * lo_mult;
    j.thr_hi = thr_cls * hi_mult;
    j.w = a.band_w[j.g];
    j.side = 2 * j.w + 1;
    j.nl = 2 * j.w + 1;
    j.boff = a.band_off[j.g];
    j.a0 = a.act_off[ad.act_base + j.d];
    j.nact = a.act_off[ad.act_base + j.d + 1] - j.a0;
    j.ndepth = ad.ndepth;
    j.gb[0] = j.gt & 1;
    j.gb[1] = (j.gt >> 1) & 1;
    j.gb[2] = (j.gt >> 2) & 1;
    return j;
}

// Norm table row stride (doubles) of one term: 4 components x (2W + 1)
// translations plus one pad, so that the 64 lane-terms fall on spread LDS
// banks.
__host__ __device__ __forceinline__ int nrm_stride(int nl) { return 4 * nl + 1; }

// Stage the chunk of active terms [base, base + cnt): per lane-term its block
// half-width and block offset, and its 1D norms [term][c][l + W] for every
// component and band translation (zero outside the term's bandwidth).
__device__ __forceinline__ void stage_terms(const KArgs &a, const AppDesc &ad, const JobCtx &j, int base, int cnt,
                                            int *ltab, long long *btab, double *nrm_s, int tid, int nth) {
    for (int e = tid; e < 64; e += nth) {
        const int m = (e < cnt) ? a.act_terms[j.a0 + base + e] : 0;
        ltab[e] = (e < cnt) ? ad.opL[m * j.ndepth + j.d] : -1;
        btab[e] = (e < cnt) ? ad.op_boff[m * j.ndepth + j.d] : 0;
    }
    for (int e = tid; e < 64 * 4 * j.nl; e += nth) {
        const int ti = e / (4 * j.nl), c = (e / j.nl) % 4, li = e % j.nl, l = li - j.w;
        double v = 0.0;
        if (ti < cnt) {
            const int m = a.act_terms[j.a0 + base + ti];
            const int L = ad.opL[m * j.ndepth + j.d];
            const int bwc = a.bw[ad.bw_off + ((long long)j.d * ad.M + m) * 4 + c];
            if (abs(l) <= bwc && abs(l) <= L)
                v = ad.op_norms[ad.op_noff[m * j.ndepth + j.d] + (long long)c * (2 * L + 1) + (l + L)];
        }
        nrm_s[ti * nrm_stride(j.nl) + c * j.nl + li] = v;
    }
}

// Screening of the (band position p, input component ft) pairs, pair index
// 8 p + ft, from pi on.  The wave first tests 64 pairs at once against the
// bound of the depth's largest term: with M the largest ((N_x N_y) N_z) of
// the active terms, fl(M ||f_ft||) is the largest of the terms' bounds
// (rounding is monotone), so a pair without a candidate flag holds no
// passing term.  For each candidate pair, in pair order, the lane-term's
// bound ((N_x N_y) N_z) ||f_ft|| is formed in FP64 in the oracle's operation
// order and a ballot appends the terms that pass the reference rule
// (bound > thr) and fall in the launch's class (bound > thr_lo and not
// bound > thr_hi).  With PAIR the job stands for the output components gt
// and gt | 4 (gt < 4), which share N_x and N_y: both bounds are formed and a
// term is appended when either passes.  emit(pos, p, ft, s, lx, ly, lz, b0,
// b1, t0, t1) is called by each appending lane with its worklist position,
// the band node's input slot s, the bounds for gt and gt | 4 and which of
// them pass (b1, t1 unused without PAIR).  Stops before a pair that could overflow the worklist (cap items)
// and leaves pi there.  Returns the item count; apps0, apps1 count the
// passing term-applications of gt and gt | 4.
template <bool PAIR, typename Emit>
__device__ __forceinline__ int screen_scan(const KArgs &a, const AppDesc &ad, const JobCtx &j, int &pi, int pi_end,
                                           int cnt, const double *nrm_s, int cap, int lane, unsigned long long &apps0,
                                           unsigned long long &apps1, Emit &&emit) {
    const bool have = lane < cnt;
    const int S = 2 * ad.mt_w + 1;
    const long long S3 = (long long)S * S * S;
    const double *mt = ad.mtab + (long long)(j.d - ad.mt_d0) * 64 * S3;
    const double *nr = nrm_s + lane * nrm_stride(j.nl);
    int n = 0;
    while (pi < pi_end) {
        const int pr = pi + lane;
        bool cand = false;
        if (pr < pi_end) {
            const int p = pr >> 3, ft = pr & 7;
            const int s = a.band[j.boff + p];
            // the NS form has no scaling-to-scaling part below the roots
            const bool ex0 = j.d > 0 && ft == 0 && j.gt == 0;
            if (s >= 0 && (PAIR || !ex0)) {
                const double fn = a.slot_norm[(long long)s * 8 + ft];
                const int lx = p % j.side - j.w, ly = (p / j.side) % j.side - j.w, lz = p / (j.side * j.side) - j.w;
                const long long li = ((long long)(lz + ad.mt_w) * S + (ly + ad.mt_w)) * S + (lx + ad.mt_w);
                const int cxy = (2 * j.gb[0] + (ft & 1)) | ((2 * j.gb[1] + ((ft >> 1) & 1)) << 2);
                const int cz = 2 * j.gb[2] + ((ft >> 2) & 1);
                double mb = ex0 ? 0.0 : mt[(cxy | (cz << 4)) * S3 + li] * fn;
                if (PAIR) mb = fmax(mb, mt[(cxy | ((cz + 2) << 4)) * S3 + li] * fn);
                cand = fn > 0.0 && mb > j.thr && mb > j.thr_lo;
            }
        }
        unsigned long long cm = __ballot(cand);
        while (cm) {
            const int k = __builtin_ctzll(cm);
            if (n + 64 > cap) {
                pi += k;
                return n;
            }
            cm &= cm - 1;
            const int p = (pi + k) >> 3, ft = (pi + k) & 7;
            const int s = a.band[j.boff + p];
            const double fn = a.slot_norm[(long long)s * 8 + ft];
            const int ll[3] = {p % j.side - j.w, (p / j.side) % j.side - j.w, p / (j.side * j.side) - j.w};
            const double nx = nr[(2 * j.gb[0] + (ft & 1)) * j.nl + (ll[0] + j.w)];
            const double ny = nr[(2 * j.gb[1] + ((ft >> 1) & 1)) * j.nl + (ll[1] + j.w)];
            const int cz = 2 * j.gb[2] + ((ft >> 2) & 1);
            const double b0 = ((nx * ny) * nr[cz * j.nl + (ll[2] + j.w)]) * fn;
            const bool ex0 = PAIR && j.d > 0 && ft == 0 && j.gt == 0;
            const bool t0 = have && !ex0 && b0 > j.thr && b0 > j.thr_lo && !(b0 > j.thr_hi);
            double b1 = 0.0;
            bool t1 = false;
            if (PAIR) {
                b1 = ((nx * ny) * nr[(cz + 2) * j.nl + (ll[2] + j.w)]) * fn;
                t1 = have && b1 > j.thr && b1 > j.thr_lo && !(b1 > j.thr_hi);
            }
            const unsigned long long mask = __ballot(t0 || t1);
            if (PAIR) {
                apps0 += __popcll(__ballot(t0));
                apps1 += __popcll(__ballot(t1));
            } else {
                apps0 += __popcll(mask);
            }
            if (t0 || t1)
                emit(n + __popcll(mask & ((1ull << lane) - 1ull)), p, ft, s, ll[0], ll[1], ll[2], b0, b1, t0, t1);
            n += __popcll(mask);
        }
        pi += 64;
    }
    return n;
}

// LDS of one job of apply_segment: the pass buffer, two operator stages,
// the worklist and the staged term table.
template <int N, typename T>
struct ApplySmem {
    T buf[JobShape<N>::BUF];
    T ops[2][JobShape<N>::OB];
    int wl[kWorkCap];
    int ltab[64];
    long long btab[64];
};

// One work segment: the terms of output job seg_job[seg] over the (band
// position, input component) pairs [seg_p0, seg_p1) in the launch's class,
// added into the FP64 output.  nrm_s
// is the norm table [term][c][l+W], 64 x nrm_stride(2 Wmax + 1) doubles.
// With PAIR the segment is a pair segment: its job stands for the output
// components gt and gt | 4 (gt < 4), whose operators differ only along z.
// An item then carries which of the two passes; its x and y passes are
// computed once and its z pass once per passing component, into two
// accumulators.  Worklist items are (p << 11) | (t1 << 10) | (t0 << 9) |
// (ft << 6) | lane-term.
template <int N, typename T, bool PAIR>
__device__ __forceinline__ void apply_segment(const KArgs &a, const T *pool, int seg, double lo_mult, double hi_mult,
                                              int cls_counter, ApplySmem<N, T> &sm, double *nrm_s) {
    using S = JobShape<N>;
    using MF = Mfma<T>;
    using acc_t = typename MF::acc_t;
    constexpr int N2 = S::N2, N3 = S::N3, NT = S::NT, KS = S::KS, TW = S::TW, NTH = S::T;
    constexpr int CP = S::CP, BUF = S::BUF;
    constexpr int OB = (PAIR ? 4 : 3) * N2, OPL = (OB + NTH - 1) / NTH;
    T *const xb = sm.buf;
    T *const yb = sm.buf + (S::SEP ? S::XSZ : 0);
    T(*ops)[S::OB] = sm.ops;
    int *wl = sm.wl;
    int *ltab = sm.ltab;
    long long *btab = sm.btab;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lc = lane & 15, lq = lane >> 4;
    for (int e = tid; e < BUF; e += NTH) xb[e] = T(0);

    const int pi_begin = a.seg_p0[seg], pi_end = a.seg_p1[seg];
    const AppDesc ad = a.apps[a.out_app[a.seg_job[seg] >> 3]];
    const JobCtx j = job_ctx(a, ad, a.seg_job[seg], lo_mult, hi_mult);
    const int g = j.g, gt = j.gt, w = j.w, side = j.side;
    const long long boff = j.boff;
    const int *gb = j.gb;
    const T *opblk = op_blocks<T>(ad);

    // this wave's tiles: t = wv*TW + tt; tiles t >= NT (N = 10: the eighth)
    // are skipped (wave-uniform), their LDS entries keep finite values that
    // only meet zero operator entries
    int ybase[TW], zbase[TW];
#pragma unroll
    for (int tt = 0; tt < TW; tt++) {
        const int col = 16 * (wv * TW + tt) + lc;
        const int q = col / N, r = col % N;
        ybase[tt] = (q * N + lq) * 16 + r;  // Xp[(jz*N + 4k + lq)][ix]
        zbase[tt] = q * CP + r + lq * N;    // Yp[iy][(4k + lq)*N + ix]
    }
    // (in FP64 only: in FP32 the branches cost more registers than the
    // padding tile's MFMAs)
    auto live = [&](int tt) { return S::NTP == NT || !std::is_same<T, double>::value || wv * TW + tt < NT; };
    acc_t acc[TW], acc1[TW];  // acc1: gt | 4 (PAIR only)
#pragma unroll
    for (int tt = 0; tt < TW; tt++) acc[tt] = acc1[tt] = acc_t{T(0), T(0), T(0), T(0)};
    unsigned long long napps = 0, napps1 = 0;
    // F tiles of band node p, component ft: x-pass A operand F[16t + lc][4k + lq]
    auto load_f = [&](int p, int ft, T (*dst)[KS]) {
        const int s = a.band[boff + p];
        const T *fsrc = pool + a.slot_off[s] + (long long)ft * N3;
#pragma unroll
        for (int tt = 0; tt < TW; tt++) {
            const int row = 16 * (wv * TW + tt) + lc;
#pragma unroll
            for (int k = 0; k < KS; k++) {
                const int jx = 4 * k + lq;
                dst[tt][k] = (row < N2 && jx < N) ? fsrc[row * N + jx] : T(0);
            }
        }
    };
    // operator blocks Ox, Oy, Oz (and Oz of gt | 4) of worklist item `it`
    // into this thread's share; z blocks of a component that does not pass
    // are left out
    auto fetch = [&](int it, T *regs) {
        const int p = it >> 11, ft = (it >> 6) & 7, b = it & 63;
        const bool t0 = !PAIR || ((it >> 9) & 1), t1 = PAIR && ((it >> 10) & 1);
        const int lx = p % side - w, ly = (p / side) % side - w, lz = p / (side * side) - w;
        const int Lb = ltab[b];
        const long long bb = btab[b];
#pragma unroll
        for (int u = 0; u < OPL; u++) {
            const int e = tid + NTH * u;
            regs[u] = T(0);
            const int dim = e / N2, r = e % N2;
            if (e < OB && (dim < 2 || (dim == 2 ? t0 : t1))) {
                // selects, not a runtime index into a small array (which would
                // live in scratch); dim 3 is z of gt | 4 (gb[2] = 1)
                const int gbd = dim == 0 ? gb[0] : (dim == 1 ? gb[1] : (dim == 2 ? gb[2] : 1));
                const int ld = dim == 0 ? lx : (dim == 1 ? ly : lz);
                const int fb = dim == 0 ? (ft & 1) : (dim == 1 ? ((ft >> 1) & 1) : ((ft >> 2) & 1));
                const int c = (gbd << 1) | fb;
                regs[u] = opblk[bb + ((long long)c * (2 * Lb + 1) + (ld + Lb)) * N2 + r];
            }
        }
    };
    auto group = [](int it) { return ((it >> 11) << 3) | ((it >> 6) & 7); };

    for (int base = 0; base < j.nact; base += 64) {
        const int cnt = min(64, j.nact - base);
        __syncthreads();
        stage_terms(a, ad, j, base, cnt, ltab, btab, nrm_s, tid, NTH);
        __syncthreads();
        int pi = pi_begin;
        while (pi < pi_end) {
            // screening into the worklist (wave 0 writes the items)
            const int n = screen_scan<PAIR>(a, ad, j, pi, pi_end, cnt, nrm_s, kWorkCap, lane, napps, napps1,
                                            [&](int pos, int p, int ft, int, int, int, int, double, double, bool t0,
                                                bool t1) {
                                                if (wv == 0)
                                                    wl[pos] = (p << 11) | ((int)t1 << 10) | ((int)t0 << 9) |
                                                              (ft << 6) | lane;
                                            });
            __syncthreads();
            if (n == 0) continue;
            // pipelined application of the worklist
            T fa[TW][KS];
            T pre[OPL];
            int it = wl[0];
            load_f(it >> 11, (it >> 6) & 7, fa);
            fetch(it, pre);
            if constexpr (S::SEP) {
                // separate buffers: item i + 1's operator blocks go into the
                // other stage at the start of item i (fetched during item
                // i - 1), so no barrier is needed before an item's operands
#pragma unroll
                for (int u = 0; u < OPL; u++) {
                    const int e = tid + NTH * u;
                    if (e < OB) ops[0][e] = pre[u];
                }
                if (n > 1) fetch(wl[1], pre);
                __syncthreads();
            }
            int stage = 0;
            for (int i = 0; i < n; i++) {
                if constexpr (!S::SEP) {
#pragma unroll
                    for (int u = 0; u < OPL; u++) {
                        const int e = tid + NTH * u;
                        if (e < OB) ops[stage][e] = pre[u];
                    }
                }
                const bool more = i + 1 < n;
                const int nx = more ? wl[i + 1] : it;
                const bool newgroup = more && group(nx) != group(it);
                const bool t0 = !PAIR || ((it >> 9) & 1), t1 = PAIR && ((it >> 10) & 1);
                if constexpr (S::SEP) {
                    if (more) {
#pragma unroll
                        for (int u = 0; u < OPL; u++) {
                            const int e = tid + NTH * u;
                            if (e < OB) ops[stage ^ 1][e] = pre[u];
                        }
                        if (i + 2 < n) fetch(wl[i + 2], pre);
                    }
                } else {
                    if (more) fetch(nx, pre);
                    // (also: the previous item's z pass has read the shared
                    // buffer)
                    __syncthreads();
                }
                // operator operands: Ox[lc][4k+lq] (B of x as Ox^T), Oy, Oz (A of y, z)
                T ox[KS], oy[KS], oz[KS], oz1[PAIR ? KS : 1];
                const T *O = ops[stage];
#pragma unroll
                for (int k = 0; k < KS; k++) {
                    const int jj = 4 * k + lq;
                    const bool v = lc < N && jj < N;
                    ox[k] = v ? O[lc * N + jj] : T(0);
                    oy[k] = v ? O[N2 + lc * N + jj] : T(0);
                    oz[k] = v ? O[2 * N2 + lc * N + jj] : T(0);
                    if (PAIR) oz1[k] = v ? O[3 * N2 + lc * N + jj] : T(0);
                }
                // x pass -> Xp[row][ix]
                {
                    acc_t dx[TW];
#pragma unroll
                    for (int tt = 0; tt < TW; tt++) dx[tt] = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
                    for (int k = 0; k < KS; k++)
#pragma unroll
                        for (int tt = 0; tt < TW; tt++)
                            if (live(tt)) dx[tt] = MF::run(fa[tt][k], ox[k], dx[tt]);
#pragma unroll
                    for (int tt = 0; tt < TW; tt++)
                        if (live(tt))
#pragma unroll
                            for (int r = 0; r < 4; r++)
                                xb[(16 * (wv * TW + tt) + MF::drow(lq, r)) * 16 + lc] = dx[tt][r];
                }
                // the F tiles are dead after the x pass: those of the next
                // item's group load during its y and z passes
                if (newgroup) load_f(nx >> 11, (nx >> 6) & 7, fa);
                __syncthreads();
                // y pass: B[jy][col] = X[jz*N + jy][ix] -> Yp[iy][col]
                {
                    acc_t dy[TW];
#pragma unroll
                    for (int tt = 0; tt < TW; tt++) dy[tt] = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
                    for (int k = 0; k < KS; k++)
#pragma unroll
                        for (int tt = 0; tt < TW; tt++)
                            if (live(tt)) dy[tt] = MF::run(oy[k], xb[ybase[tt] + 64 * k], dy[tt]);
                    // one shared buffer: every wave has read X before Y
                    // overwrites it
                    if constexpr (!S::SEP) __syncthreads();
#pragma unroll
                    for (int tt = 0; tt < TW; tt++)
                        if (live(tt))
#pragma unroll
                            for (int r = 0; r < 4; r++)
                                yb[MF::drow(lq, r) * CP + 16 * (wv * TW + tt) + lc] = dy[tt][r];
                }
                __syncthreads();
                // z pass: B[jz][col'] = Yp[iy][jz*N + ix], col' = iy*N + ix,
                // once per passing component (item-uniform branches)
                if (!PAIR || (t0 && t1)) {
#pragma unroll
                    for (int k = 0; k < KS; k++)
#pragma unroll
                        for (int tt = 0; tt < TW; tt++)
                            if (live(tt)) {
                                const T yv = yb[zbase[tt] + 4 * k * N];
                                acc[tt] = MF::run(oz[k], yv, acc[tt]);
                                if (PAIR) acc1[tt] = MF::run(oz1[k], yv, acc1[tt]);
                            }
                } else if (t0) {
#pragma unroll
                    for (int k = 0; k < KS; k++)
#pragma unroll
                        for (int tt = 0; tt < TW; tt++)
                            if (live(tt)) acc[tt] = MF::run(oz[k], yb[zbase[tt] + 4 * k * N], acc[tt]);
                } else {
#pragma unroll
                    for (int k = 0; k < KS; k++)
#pragma unroll
                        for (int tt = 0; tt < TW; tt++)
                            if (live(tt)) acc1[tt] = MF::run(oz1[k], yb[zbase[tt] + 4 * k * N], acc1[tt]);
                }
                it = nx;
                stage ^= 1;
            }
            __syncthreads();
        }
    }
    if (tid == 0 && napps + napps1) {
        atomicAdd(&a.counters[0], napps + napps1);
        if (cls_counter > 0) atomicAdd(&a.counters[cls_counter], napps + napps1);
    }
    // acc[tt][r] = out[iz = drow(lq, r)][col = 16t + lc], added in FP64 to the
    // (zeroed) output: several segments and launches may share a job
    auto store = [&](const acc_t *ac, int comp) {
        double *dst = a.out_ns + (long long)g * 8 * N3 + (long long)comp * N3;
#pragma unroll
        for (int tt = 0; tt < TW; tt++) {
            const int col = 16 * (wv * TW + tt) + lc;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int iz = MF::drow(lq, r);
                if (col < N2 && iz < N) atomicAdd(&dst[iz * N2 + col], (double)ac[tt][r]);
            }
        }
    };
    if (napps) store(acc, gt);
    if (PAIR && napps1) store(acc1, gt | 4);
}

// ------------------------------------------------------ FP16 split products
//
// Each operand value v of a block, scaled into (-1, 1) by the block's
// power-of-two scale, is split as v = hi + lo (both FP16, lo = v - hi
// rounded).  A product is formed as hi*hi + lo*hi + hi*lo (lo*lo is below the
// representation error) by one v_mfma_f32_16x16x32_f16 per 16x16 tile: lane
// chunk c = lane / 16 holds K slots 8c..8c+7, the data side supplies
//   [hi 0..7 | lo 0..7 | hi 0..7 | tail]
// and the operator side
//   [hi 0..7 | hi 0..7 | lo 0..7 | tail]
// where the tails carry the same three products for j = 8, 9 (N = 9, 10:
// data {hi8 hi9 hi8 hi9 lo8 lo9 lo8 lo9}, operator {hi8 hi9 lo8 lo9 hi8 hi9 0 0};
// the data tail is the plain {hi | lo} of lines 8, 9, 8, 9, so a pass that
// computes lines 10, 11 as copies of lines 8, 9 stores it like any chunk)
// or hi*hi + lo*hi for j = 8..11 (N = 11, 12: data {hi8..11 lo8..11},
// operator {hi8..11 hi8..11}); for N = 11, 12 an extra chunk holds hi8..11
// (data) / lo8..11 (operator) for a v_mfma_f32_16x16x16f16 in lane chunk 0.
// A row stores the distinct chunks once, [hi 0..7 | lo 0..7 | tail] (+ the
// extra chunk and a pad chunk for N > 10), and each lane reads the chunk its
// K slots need with one 16-byte load: data lanes chunks (0, 1, 0, 2),
// operator lanes chunks (0, 0, 1, 2).
typedef _Float16 mwg_h8 __attribute__((ext_vector_type(8)));
typedef _Float16 mwg_h4 __attribute__((ext_vector_type(4)));

#ifndef MWG_SPLIT_MASK_ALL
#define MWG_SPLIT_MASK_ALL 0
#endif
template <int N>
struct SplitShape {
    static constexpr int N2 = N * N, N3 = N2 * N;
    static constexpr int G = (N + 3) / 4;       // 4-line groups per line of N
    static constexpr int NG = N * G;            // groups of an (outer, inner) plane
    static constexpr int NTG = (NG + 3) / 4;    // 16-row tiles of groups (x and y passes)
    static constexpr int NTZ = (N2 + 15) / 16;  // 16-column tiles (z pass)
    // waves per job: as many as keep every wave's tiles busy (N <= 4: one
    // tile per pass; N = 5, 6 and 11, 12: tile counts of 3 and 9)
    static constexpr int NW = N <= 4 ? 1 : (N <= 6 || N > 10 ? 3 : 4);
    static constexpr int T = 64 * NW;
    static constexpr int TWG = (NTG + NW - 1) / NW;
    static constexpr int TWZ = (NTZ + NW - 1) / NW;
    static constexpr bool TWO = N > 10;
    // waves per SIMD the register allocation is held to (VGPRs <= 512 / WPE):
    // N > 10 at 3 (64 Helmholtz applies at k = 11: -6 % although 140 bytes
    // per lane spill); no limit below, where a cap of 4 spills at N = 5, 7, 8
    // or schedules worse
    static constexpr int WPE = N > 10 ? 3 : 1;
    // N <= 8: every phase issues its six loads unconditionally (exact wait
    // counts; 136 exchange pair applies at k = 7: -25 %); above, the extra
    // live registers cost occupancy or spill and the loads stay conditional
    static constexpr bool UNC = N <= 8;
    // N <= 10: screening resolves each item's F block (address and scale), so
    // a phase loads F without a chain of dependent loads; at N > 10 the extra
    // screening work costs more than it saves (64 applies at k = 11: +6 %)
    static constexpr bool CB = N <= 10;
    // lanes without a result row store nothing (else they write a scratch row)
    static constexpr bool MASK = MWG_SPLIT_MASK_ALL || N <= 8;
    // 16-byte chunks per row: hi, lo, tail (+ the extra chunk for N > 10).
    // For N <= 10 the odd number of chunks makes the row pitch spread 16
    // consecutive rows over distinct LDS banks; for N > 10 the four chunks
    // keep the double-buffered intermediates within 64 KB of LDS.
    static constexpr int NC = TWO ? 4 : 3;
    static constexpr int RH = 8 * NC;           // halfs per global row
    // LDS rows of the intermediates: hi, lo, tail in 48 bytes (an odd number
    // of 16-byte chunks, so 16-byte reads of 8 consecutive rows hit distinct
    // banks); the extra chunk of N > 10 sits in a region of its own, 8 bytes
    // per row, after the rows.
    static constexpr int RL = 24;               // halfs per LDS row
    static constexpr int RE = TWO ? 4 : 0;      // halfs per LDS extra chunk
    static constexpr int EXT0 = (N2 + 2) * RL;          // start of the extra chunks
    static constexpr int BUFH = (N2 + 2) * (RL + RE);   // halfs per intermediate buffer
};

// Split row of n <= 12 values u (|u| < 1): data side (DATA) or operator side.
template <int N, bool DATA>
__device__ __forceinline__ void split_row(const double *u, _Float16 *row) {
    _Float16 hi[12], lo[12];
#pragma unroll
    for (int j = 0; j < 12; j++) {
        const double v = j < N ? u[j] : 0.0;
        hi[j] = (_Float16)v;
        lo[j] = (_Float16)(v - (double)hi[j]);
    }
    const _Float16 z = (_Float16)0.0f;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        row[j] = hi[j];
        row[8 + j] = lo[j];
    }
    _Float16 t[8];
    if (N <= 8) {
        for (int e = 0; e < 8; e++) t[e] = z;
    } else if (N <= 10) {
        const _Float16 td[8] = {hi[8], hi[9], hi[8], hi[9], lo[8], lo[9], lo[8], lo[9]};
        const _Float16 to[8] = {hi[8], hi[9], lo[8], lo[9], hi[8], hi[9], z, z};
        for (int e = 0; e < 8; e++) t[e] = DATA ? td[e] : to[e];
    } else {
        for (int e = 0; e < 4; e++) {
            t[e] = hi[8 + e];
            t[4 + e] = DATA ? lo[8 + e] : hi[8 + e];
        }
    }
#pragma unroll
    for (int e = 0; e < 8; e++) row[16 + e] = t[e];
    if (SplitShape<N>::TWO) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            row[24 + e] = DATA ? hi[8 + e] : lo[8 + e];
            row[28 + e] = z;
        }
    }
}

// The second MFMA gets its own zero accumulator: issued directly on the
// accumulator just written by the 16x16x32 MFMA (no wait states between
// them in the generated code) it gave wrong sums on gfx950, while separate
// accumulators added afterwards are exact.
template <int N>
__device__ __forceinline__ mwg_float4 split_mfma(const mwg_h8 &a, const mwg_h4 &ae, const mwg_h8 &b, const mwg_h4 &be) {
    mwg_float4 c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, mwg_float4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    if (SplitShape<N>::TWO) {
        const mwg_float4 e = __builtin_amdgcn_mfma_f32_16x16x16f16(ae, be, mwg_float4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; r++) c[r] += e[r];
    }
    return c;
}

// a for P = 0, b for P = 1, chosen at compile time
template <int P, typename T>
__device__ __forceinline__ T &pick(T &a, T &b) {
    if constexpr (P == 0)
        return a;
    else
        return b;
}

// Global-memory view of a pointer, so that loads through it are global, not
// flat, loads.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gmem(const T *p) {
    return (const __attribute__((address_space(1))) T *)p;
}

// Worklist item of the split kernel: the operator block ids (block = N x N
// values of one term, depth, component and translation) of x, y and of z
// for the output components gt and gt | 4, and the input group (band
// position << 3 | component) with flag bits: kZ0Bit, kZ1Bit when the term
// passes for gt, gt | 4; kSingleBit when the item's intermediates are
// rounded once (no lo part).
struct SplitItem {
    int bx, by, bz0, bz1, grp, cb;
};
constexpr int kZ0Bit = 28, kZ1Bit = 29, kSingleBit = 30;

// Per-item passes with FP16 split operands.  The screening rule is that of
// k_apply; a job is an output node g and a pair of output components
// (gt, gt | 4), gt < 4, whose operators differ only along z, so an item
// (band node, ft, term) runs the x and y passes once and the z pass for
// each component of the pair its term passes for.  The screening lane (one
// term) also resolves the item's operator block ids.  With unit-scaled
// operands F_u, O_u:
//   x: X[(jz,jy)][ix] = F_u . Ox_u^T, stored transposed as Xt[(jz,ix)][jy]
//   y: Y[(ix,jz)][iy] = Xt . Oy_u^T,  stored as Yt[(ix,iy)][jz]
//   z: Z[iz][(ix,iy)] = Oz_u . Yt^T,  acc_h += sF sOx sOy sOz_h Z  (FP32)
//      for h = 0, 1 (gt, gt | 4), one read of Yt for both
// The rows of the x and y passes are taken in groups of four consecutive
// inner indices (x: (jz, 4 jy), y: (ix, 4 jz), the y groups ordered with ix
// fastest), so every lane stores its four results with aligned 8-byte
// writes, the 16 lanes of a store write 16 consecutive rows and the 8 lanes
// of a 16-byte read hit distinct LDS banks.  Intermediates are split again (hi: the value
// cut to an 11-bit mantissa, lo: the exact FP32 remainder, both packed to
// FP16) before the next pass; their magnitudes stay below N^2.  Items whose
// bound is at most single_mult * thr_cls keep the intermediates rounded
// once (lo = 0): their arithmetic error stays below ~1e-3 of the bound.
// Lanes whose operand row lies past the last line read a line that exists
// (line j - 2 or N - 1, or row 0 of the block): their products meet zero
// operand slots or land in output rows and columns that are not stored.
// Lanes without a row to write store nothing at N <= 8 (an exec mask: all
// of them writing one scratch row serialises the store), a scratch row above,
// and the single-rounding choice is a bit mask, so no pass branches on the
// item.
constexpr int kSplitCap = 384;  // worklist items per batch (>= 64 + 128)

// LDS of one job of split_segment: the double-buffered intermediates, the
// worklist with its scale products and the staged term table.
template <int N>
struct SplitSmem {
    alignas(16) _Float16 xt[2 * SplitShape<N>::BUFH];  // X of items q, q - 1 (by parity)
    alignas(16) _Float16 yt[2 * SplitShape<N>::BUFH];  // Y of items q - 1, q - 2
    SplitItem wl[kSplitCap];
    float2 wsig[kSplitCap];  // products of the item's operator scales, z of gt and of gt | 4
    int ltab[64];
    long long btab[64];
};

// One pair segment with FP16 split products (the passes described above);
// nrm_s as for apply_segment.
template <int N>
__device__ __forceinline__ void split_segment(const KArgs &a, const _Float16 *pool_h, const double *pool_sc, int seg,
                                              double lo_mult, double hi_mult, double single_mult, int cls_counter,
                                              SplitSmem<N> &sm, double *nrm_s) {
    using S = SplitShape<N>;
    constexpr int N2 = S::N2, N3 = S::N3, G = S::G, NG = S::NG, NTZ = S::NTZ;
    constexpr int TWG = S::TWG, TWZ = S::TWZ, NTH = S::T, RH = S::RH, RL = S::RL, RE = S::RE, NW = S::NW;
    constexpr int ZROW = N2, DROW = N2 + 1;  // all-zero row, scratch row
    constexpr int EXT0 = S::EXT0;              // start of the extra chunks
    constexpr int BUFH = S::BUFH;              // halfs per intermediate buffer
    _Float16 *xt = sm.xt, *yt = sm.yt;
    SplitItem *wl = sm.wl;
    float2 *wsig = sm.wsig;
    int *ltab = sm.ltab;
    long long *btab = sm.btab;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lc = lane & 15, lq = lane >> 4;

    const int pi_begin = a.seg_p0[seg], pi_end = a.seg_p1[seg];
    const AppDesc ad = a.apps[a.out_app[a.seg_job[seg] >> 3]];
    const JobCtx j = job_ctx(a, ad, a.seg_job[seg], lo_mult, hi_mult);
    const double thr_single = (j.thr > ad.thr_floor ? j.thr : ad.thr_floor) * single_mult;
    const _Float16 z16 = (_Float16)0.0f;
    const mwg_h8 zero8 = {z16, z16, z16, z16, z16, z16, z16, z16};
    const mwg_h4 zero4 = {z16, z16, z16, z16};
    // rows and chunks the passes never write are read as operands: zeros
    for (int e = tid; e < 2 * BUFH / 8; e += NTH) {
        reinterpret_cast<mwg_h8 *>(xt)[e] = zero8;
        reinterpret_cast<mwg_h8 *>(yt)[e] = zero8;
    }
    const __attribute__((address_space(1))) _Float16 *opg = gmem(ad.blk_h);
    const __attribute__((address_space(1))) double *opsc = gmem(ad.blk_hsc);
    const __attribute__((address_space(1))) _Float16 *poolg = gmem(pool_h);

    // chunk offsets (halfs) this lane reads: data side (0, 1, 0, 2), operator
    // side (0, 0, 1, 2); the extra chunk (N > 10) only in lane chunk 0
    const int dch = 8 * (lq == 1 ? 1 : (lq == 3 ? 2 : 0));
    const int och = 8 * (lq == 2 ? 1 : (lq == 3 ? 2 : 0));
    // line read for a row index past the last line
    auto line = [](int jj) { return jj < N ? jj : (jj - 2 < N ? jj - 2 : N - 1); };

    // Per-lane offsets (halfs), fixed for the kernel: LDS stores of the x and
    // y passes (hi part, lo part, extra chunk), LDS operand reads of the y and
    // z passes (chunk, extra chunk), F-block reads of the x pass (chunk,
    // extra chunk) and operator-block reads (row min(lc, N - 1)).
    int xs1[TWG], xs2[TWG], xse[TWG], ys1[TWG], ys2[TWG], yse[TWG];
    int yl[TWG], yle[TWG], fl[TWG], fle[TWG], zl[TWZ], zle[TWZ];
    // lanes with a result row to store, and those whose row is a tail row
    // (the extra chunk, N > 10): the others store nothing (no two lanes of a
    // store meet at one scratch address)
    bool sok[TWG], xtl[TWG], ytl[TWG];
#pragma unroll
    for (int tt = 0; tt < TWG; tt++) {
        const int t = wv + NW * tt;
        const int g4 = 4 * t + lq;  // group of this lane's four results
        const bool okx = g4 < NG && lc < N;
        sok[tt] = okx;
        // x: group g4 = (jz, jy / 4) = (g4 / G, g4 % G), lane ix = lc -> Xt row jz N + ix
        {
            const int ri = okx ? (g4 / G) * N + lc : DROW, row = ri * RL, j0 = 4 * (g4 % G);
            const bool tail = N > 8 && j0 >= 8;
            xtl[tt] = tail;
            xs1[tt] = row + (tail ? 16 : j0);
            xs2[tt] = row + (tail ? 20 : 8 + j0);
            xse[tt] = EXT0 + (tail ? ri : DROW) * RE;
        }
        // y: group g4 = (ix, jz / 4) = (g4 % N, g4 / N), lane iy = lc -> Yt row ix N + iy
        {
            const int ri = okx ? (g4 % N) * N + lc : DROW, row = ri * RL, j0 = 4 * (g4 / N);
            const bool tail = N > 8 && j0 >= 8;
            ytl[tt] = tail;
            ys1[tt] = row + (tail ? 16 : j0);
            ys2[tt] = row + (tail ? 20 : 8 + j0);
            yse[tt] = EXT0 + (tail ? ri : DROW) * RE;
        }
        // operand rows of lane lc: group g5 = 4 t + lc / 4, line 4 (g5 % G or
        // g5 / N) + lc % 4 of the x pass (F: jz = g5 / G) and the y pass
        // (Xt: ix = g5 % N)
        const int g5 = 4 * t + (lc >> 2);
        {
            const int jz = 4 * (g5 / N) + (lc & 3);
            const int ri = g5 < NG ? line(jz) * N + g5 % N : ZROW;
            yl[tt] = ri * RL + dch;
            yle[tt] = EXT0 + (lq == 0 ? ri : ZROW) * RE;
        }
        {
            const int jy = 4 * (g5 % G) + (lc & 3);
            const int r5 = (g5 < NG ? (g5 / G) * N + line(jy) : 0) * RH;
            fl[tt] = r5 + dch;
            fle[tt] = r5 + 24;
        }
    }
#pragma unroll
    for (int tt = 0; tt < TWZ; tt++) {
        const int col = 16 * (wv + NW * tt) + lc;
        const int ri = col < N2 ? col : ZROW;
        zl[tt] = ri * RL + dch;
        zle[tt] = EXT0 + (lq == 0 ? ri : ZROW) * RE;
    }
    const int ol = min(lc, N - 1) * RH + och, ole = min(lc, N - 1) * RH + 24;

    mwg_float4 acc0[TWZ], acc1[TWZ];  // Z sums of gt and gt | 4
#pragma unroll
    for (int tt = 0; tt < TWZ; tt++) acc0[tt] = acc1[tt] = mwg_float4{0.f, 0.f, 0.f, 0.f};
    unsigned long long napps0 = 0, napps1 = 0;

    // four results of one tile into a data-side split row: hi keeps the bits
    // of msk (0xffffe000: an 11-bit mantissa, so hi converts exactly and
    // lo = v - hi is the exact remainder; all ones: single rounding, lo = 0)
    auto store4 = [&](_Float16 *base, int o1, int o2, int oe, const mwg_float4 &v, unsigned msk, bool ok, bool ext) {
        float h[4], l[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            h[r] = __uint_as_float(__float_as_uint(v[r]) & msk);
            l[r] = v[r] - h[r];
        }
        const mwg_h4 hi = mwg_h4{(_Float16)h[0], (_Float16)h[1], (_Float16)h[2], (_Float16)h[3]};
        const mwg_h4 lo = mwg_h4{(_Float16)l[0], (_Float16)l[1], (_Float16)l[2], (_Float16)l[3]};
        if (!S::MASK || ok) {
            *reinterpret_cast<mwg_h4 *>(base + o1) = hi;
            *reinterpret_cast<mwg_h4 *>(base + o2) = lo;
            if (S::TWO && (!S::MASK || ext)) *reinterpret_cast<mwg_h4 *>(base + oe) = hi;
        }
    };
    auto lds_operand = [&](const _Float16 *base, int off, int offe, mwg_h8 &v, mwg_h4 &e) {
        v = *reinterpret_cast<const mwg_h8 *>(base + off);
        if (S::TWO) e = *reinterpret_cast<const mwg_h4 *>(base + offe);
    };

    // x-pass A operand (data side): rows (jz, jy) of the F block, in groups
    mwg_h8 fa[TWG];
    mwg_h4 fe[TWG];
#pragma unroll
    for (int tt = 0; tt < TWG; tt++) fe[tt] = zero4;
    // the F rows of the x pass: component block cb of the pool, or (without
    // CB) that of group grp = band position << 3 | ft, with its scale sF
    float sF = 1.f;
    auto load_f = [&](int cbg) {
        long long cb = cbg;
        if constexpr (!S::CB) {
            const int s = a.band[j.boff + (cbg >> 3)];
            cb = a.slot_off[s] / N3 + (cbg & 7);
            sF = (float)pool_sc[cb];
        }
        const __attribute__((address_space(1))) _Float16 *fb = poolg + cb * (N2 * RH);
#pragma unroll
        for (int tt = 0; tt < TWG; tt++) {
            fa[tt] = *reinterpret_cast<const __attribute__((address_space(1))) mwg_h8 *>(fb + fl[tt]);
            if (S::TWO) {
                const mwg_h4 e = *reinterpret_cast<const __attribute__((address_space(1))) mwg_h4 *>(fb + fle[tt]);
                fe[tt] = lq == 0 ? e : zero4;
            }
        }
    };
    // operator block blk (uniform): row min(lc, N - 1), the lane's chunk; the
    // extra chunk of lanes lq > 0 meets zero data slots
    auto load_op = [&](int blk, mwg_h8 &v, mwg_h4 &e) {
        const __attribute__((address_space(1))) _Float16 *b = opg + (long long)blk * (N * RH);
        v = *reinterpret_cast<const __attribute__((address_space(1))) mwg_h8 *>(b + ol);
        if (S::TWO) e = *reinterpret_cast<const __attribute__((address_space(1))) mwg_h4 *>(b + ole);
    };
    // The passes of one phase: x of item q into X[q & 1], y of item q - 1
    // from X[(q - 1) & 1] into Y[(q - 1) & 1], z of item q - 2 from
    // Y[q & 1] into acc.  Each buffer a pass reads was written in the
    // previous phase and each buffer it writes was last read in the previous
    // phase, so one barrier per phase orders them.
    auto x_pass = [&](_Float16 *X, const mwg_h8 &o, const mwg_h4 &oe, unsigned msk) {
        mwg_float4 d[TWG];
#pragma unroll
        for (int tt = 0; tt < TWG; tt++) d[tt] = split_mfma<N>(fa[tt], fe[tt], o, oe);
#pragma unroll
        for (int tt = 0; tt < TWG; tt++) store4(X, xs1[tt], xs2[tt], xse[tt], d[tt], msk, sok[tt], xtl[tt]);
    };
    auto y_pass = [&](const _Float16 *X, _Float16 *Y, const mwg_h8 &o, const mwg_h4 &oe, unsigned msk) {
        mwg_h8 va[TWG];
        mwg_h4 ve[TWG];
#pragma unroll
        for (int tt = 0; tt < TWG; tt++) {
            ve[tt] = zero4;
            lds_operand(X, yl[tt], yle[tt], va[tt], ve[tt]);
        }
        mwg_float4 d[TWG];
#pragma unroll
        for (int tt = 0; tt < TWG; tt++) d[tt] = split_mfma<N>(va[tt], ve[tt], o, oe);
#pragma unroll
        for (int tt = 0; tt < TWG; tt++) store4(Y, ys1[tt], ys2[tt], yse[tt], d[tt], msk, sok[tt], ytl[tt]);
    };
    // z of gt (zm bit 0, operator o0) and of gt | 4 (bit 1, operator o1)
    auto z_pass = [&](const _Float16 *Y, const mwg_h8 &o0, const mwg_h4 &oe0, const mwg_h8 &o1, const mwg_h4 &oe1,
                      float2 sigma, int zm) {
        mwg_h8 vb[TWZ];
        mwg_h4 vbe[TWZ];
#pragma unroll
        for (int tt = 0; tt < TWZ; tt++) {
            vbe[tt] = zero4;
            lds_operand(Y, zl[tt], zle[tt], vb[tt], vbe[tt]);
        }
        if (zm & 1) {
#pragma unroll
            for (int tt = 0; tt < TWZ; tt++) {
                const mwg_float4 dz = split_mfma<N>(o0, oe0, vb[tt], vbe[tt]);
#pragma unroll
                for (int r = 0; r < 4; r++) acc0[tt][r] = fmaf(sigma.x, dz[r], acc0[tt][r]);
            }
        }
        if (zm & 2) {
#pragma unroll
            for (int tt = 0; tt < TWZ; tt++) {
                const mwg_float4 dz = split_mfma<N>(o1, oe1, vb[tt], vbe[tt]);
#pragma unroll
                for (int r = 0; r < 4; r++) acc1[tt][r] = fmaf(sigma.y, dz[r], acc1[tt][r]);
            }
        }
    };

    for (int base = 0; base < j.nact; base += 64) {
        const int cnt = min(64, j.nact - base);
        __syncthreads();
        stage_terms(a, ad, j, base, cnt, ltab, btab, nrm_s, tid, NTH);
        __syncthreads();
        // this lane's term: operator block base and half-width
        const int my_L = ltab[lane];
        const int my_b0 = (int)(btab[lane] / N2) + my_L;
        int pi = pi_begin;
        while (pi < pi_end) {
            // screening into SplitItems: the lane-term also resolves the
            // item's three operator block ids; wave 0 writes the items
            const int n = screen_scan<true>(
                a, ad, j, pi, pi_end, cnt, nrm_s, kSplitCap, lane, napps0, napps1,
                [&](int pos, int p, int ft, int slot, int lx, int ly, int lz, double b0, double b1, bool t0, bool t1) {
                    if (wv != 0) return;
                    // a passing term has |l| <= L in every direction, so all
                    // four block ids exist
                    const int span = 2 * my_L + 1;
                    SplitItem itm;
                    itm.bx = my_b0 + (((j.gb[0] << 1) | (ft & 1)) * span) + lx;
                    itm.by = my_b0 + (((j.gb[1] << 1) | ((ft >> 1) & 1)) * span) + ly;
                    itm.bz0 = my_b0 + (((ft >> 2) & 1) * span) + lz;
                    itm.bz1 = itm.bz0 + 2 * span;
                    const double bmax = fmax(t0 ? b0 : 0.0, t1 ? b1 : 0.0);
                    itm.grp = (p << 3) | ft | ((bmax <= thr_single ? 1 : 0) << kSingleBit) | ((t0 ? 1 : 0) << kZ0Bit) |
                              ((t1 ? 1 : 0) << kZ1Bit);
                    // the input component block (its F rows in the split
                    // pool) and its scale, resolved here so that a phase
                    // starts the F load without a chain of dependent loads
                    double sxy = opsc[itm.bx] * opsc[itm.by];
                    if constexpr (SplitShape<N>::CB) {
                        itm.cb = (int)(a.slot_off[slot] / N3) + ft;
                        sxy *= pool_sc[itm.cb];
                    }
                    wl[pos] = itm;
                    wsig[pos] = make_float2(t0 ? (float)(sxy * opsc[itm.bz0]) : 0.f,
                                            t1 ? (float)(sxy * opsc[itm.bz1]) : 0.f);
                });
            __syncthreads();
            if (n == 0) continue;
            // software pipeline over the batch: phase q runs x(q), y(q - 1),
            // z(q - 2) with operator set q & 1 while the next phase's
            // operators load into the other set; phases are unrolled in pairs
            // so that the set and the LDS buffers are fixed in each copy
            const int mask_grp = (1 << kZ0Bit) - 1;
            int cur_grp = -1;  // the group whose F rows are loaded (N > 8)
            mwg_h8 o0[4], o1[4];  // operator sets: Ox, Oy, Oz of gt, Oz of gt | 4
            mwg_h4 e0[4], e1[4];
#pragma unroll
            for (int u = 0; u < 4; u++) e0[u] = e1[u] = zero4;
            float2 sig1 = make_float2(0.f, 0.f), sig2 = make_float2(0.f, 0.f);  // scales of items q - 1, q - 2
            int zm1 = 0, zm2 = 0;  // z masks of items q - 1, q - 2
            unsigned msk1 = 0u;    // hi mask of item q - 1
            load_op(__builtin_amdgcn_readfirstlane(wl[0].bx), o0[0], e0[0]);
            // full: a steady phase (all three passes), known at compile time
            auto phase = [&](int q, auto par, auto full) {
                constexpr int P = decltype(par)::value;
                constexpr bool FULL = decltype(full)::value;
                auto &oc = pick<P>(o0, o1);
                auto &ec = pick<P>(e0, e1);
                auto &on = pick<1 - P>(o0, o1);
                auto &en = pick<1 - P>(e0, e1);
                const bool hx = FULL || q < n, hy = FULL || (q >= 1 && q <= n), hz = FULL || q >= 2;
                float2 sig0 = make_float2(0.f, 0.f);
                unsigned msk0 = 0u;
                int zm0 = 0;
                if constexpr (S::UNC) {
                    // The F rows of item q first, then the next phase's
                    // operators Ox(q + 1), Oy(q), Oz(q - 1) of gt and gt | 4.
                    // All six loads are issued in every phase, their items
                    // clamped into the batch (a repeated F block or an unused
                    // operator block is a cache hit), so the loads complete
                    // in a fixed order and each pass waits only for its own
                    // operands: y and z for the previous phase's loads, x
                    // (run last) for F.
                    const int qx = q < n ? q : n - 1, qn = q + 1 < n ? q + 1 : n - 1;
                    const int qz = q >= 1 && q <= n ? q - 1 : 0;
                    const int g = __builtin_amdgcn_readfirstlane(wl[qx].grp);
                    load_f(__builtin_amdgcn_readfirstlane(wl[qx].cb));
                    load_op(__builtin_amdgcn_readfirstlane(wl[qn].bx), on[0], en[0]);
                    load_op(__builtin_amdgcn_readfirstlane(wl[qx].by), on[1], en[1]);
                    load_op(__builtin_amdgcn_readfirstlane(wl[qz].bz0), on[2], en[2]);
                    load_op(__builtin_amdgcn_readfirstlane(wl[qz].bz1), on[3], en[3]);
                    if (hx) {
                        msk0 = ((g >> kSingleBit) & 1) ? 0xffffffffu : 0xffffe000u;
                        zm0 = (g >> kZ0Bit) & 3;
                        sig0 = wsig[q];
                    }
                    if (hy) y_pass(xt + (1 - P) * BUFH, yt + (1 - P) * BUFH, oc[1], ec[1], msk1);
                    if (hz) z_pass(yt + P * BUFH, oc[2], ec[2], oc[3], ec[3], sig2, zm2);
                    if (hx) x_pass(xt + P * BUFH, oc[0], ec[0], msk0);
                } else {
                    // next phase's operators: Ox(q + 1), Oy(q), Oz(q - 1) of
                    // the components item q - 1 passes for; the F rows of
                    // item q only when its group changes (registers are
                    // short at N > 8)
                    if (q + 1 < n) load_op(__builtin_amdgcn_readfirstlane(wl[q + 1].bx), on[0], en[0]);
                    if (hx) load_op(__builtin_amdgcn_readfirstlane(wl[q].by), on[1], en[1]);
                    if (hy && (zm1 & 1)) load_op(__builtin_amdgcn_readfirstlane(wl[q - 1].bz0), on[2], en[2]);
                    if (hy && (zm1 & 2)) load_op(__builtin_amdgcn_readfirstlane(wl[q - 1].bz1), on[3], en[3]);
                    if (hx) {
                        const int g = __builtin_amdgcn_readfirstlane(wl[q].grp);
                        if ((g & mask_grp) != cur_grp) {
                            cur_grp = g & mask_grp;
                            load_f(S::CB ? __builtin_amdgcn_readfirstlane(wl[q].cb) : cur_grp);
                        }
                        msk0 = ((g >> kSingleBit) & 1) ? 0xffffffffu : 0xffffe000u;
                        zm0 = (g >> kZ0Bit) & 3;
                        const float2 w2 = wsig[q];
                        sig0 = make_float2(sF * w2.x, sF * w2.y);
                        x_pass(xt + P * BUFH, oc[0], ec[0], msk0);
                    }
                    if (hy) y_pass(xt + (1 - P) * BUFH, yt + (1 - P) * BUFH, oc[1], ec[1], msk1);
                    if (hz) z_pass(yt + P * BUFH, oc[2], ec[2], oc[3], ec[3], sig2, zm2);
                }
                __syncthreads();
                sig2 = sig1;
                sig1 = sig0;
                zm2 = zm1;
                zm1 = zm0;
                msk1 = msk0;
            };
            // N <= 8 (registers to spare): steady phase pairs get a copy of
            // their own without the per-pass conditions (136 exchange pair
            // applies at k = 7: -13 %; at N >= 9 the copy costs a wave per
            // SIMD and is slower)
            for (int q = 0; q < n + 2; q += 2) {
                if (N <= 8 && q >= 2 && q + 1 < n) {
                    phase(q, std::integral_constant<int, 0>{}, std::integral_constant<bool, N <= 8>{});
                    phase(q + 1, std::integral_constant<int, 1>{}, std::integral_constant<bool, N <= 8>{});
                } else {
                    phase(q, std::integral_constant<int, 0>{}, std::false_type{});
                    if (q + 1 < n + 2) phase(q + 1, std::integral_constant<int, 1>{}, std::false_type{});
                }
            }
            __syncthreads();
        }
    }
    const unsigned long long napps = napps0 + napps1;
    if (tid == 0 && napps) {
        atomicAdd(&a.counters[0], napps);
        if (cls_counter > 0) atomicAdd(&a.counters[cls_counter], napps);
    }
    // acc[tt][r] = Z[iz = 4 lq + r][col = 16 t + lc] with col = ix N + iy:
    // put in (iz, iy, ix) order in LDS, then added in FP64 to the output
    // with consecutive lanes on consecutive elements
    static_assert(N3 * sizeof(float) <= sizeof(sm.xt), "output staging does not fit");
    float *zs = reinterpret_cast<float *>(xt);
    auto write_out = [&](const mwg_float4 (&acc)[TWZ], int gt) {
        __syncthreads();
#pragma unroll
        for (int tt = 0; tt < TWZ; tt++) {
            const int t = wv + NW * tt;
            const int col = 16 * t + lc;
            if (t >= NTZ || col >= N2) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int iz = 4 * lq + r;
                if (iz < N) zs[iz * N2 + (col % N) * N + col / N] = acc[tt][r];
            }
        }
        __syncthreads();
        double *dst = a.out_ns + (long long)j.g * 8 * N3 + (long long)gt * N3;
        for (int e = tid; e < N3; e += NTH) atomicAdd(&dst[e], (double)zs[e]);
    };
    if (napps0) write_out(acc0, j.gt);
    if (napps1) write_out(acc1, j.gt | 4);
}

// ------------------------------------------------------ persistent launches
//
// Each launch is a persistent grid over a queue of work entries: a workgroup
// takes the next entry from an atomic head (zeroed before the launch), runs
// it and takes another until the queue is empty, so every workgroup reaches
// the exit.  The body's LDS is a static struct (compile-time LDS addresses);
// the norm table is the dynamic LDS.

// Next queue entry for the workgroup (the same value in every thread).
__device__ __forceinline__ int next_entry(unsigned long long *head, int *s_next) {
    // every thread is done with the previous entry (its LDS) before any
    // thread starts the next one
    __syncthreads();
    if (threadIdx.x == 0) *s_next = (int)atomicAdd(head, 1ull);
    __syncthreads();
    // uniform: the job context derived from it stays in scalar registers
    return __builtin_amdgcn_readfirstlane(*s_next);
}

// FP64 or FP32 per-node calculator over segments seg_list[0, n) (or segments
// seg0 .. seg0 + n - 1 without a list) in the class (lo_mult thr_cls,
// hi_mult thr_cls]; PAIR: the segments are pair segments.  Held to two waves
// per SIMD: the FP64 pair jobs at N = 7, 10, 11 would otherwise take more
// than 256 registers and run one wave per SIMD (N = 10: 8.4 ms against
// 5.8 ms with 72 bytes per lane spilled, 116 Helmholtz terms at k = 9).
template <int N, typename T, bool PAIR>
__global__ __launch_bounds__(JobShape<N>::T) __attribute__((amdgpu_waves_per_eu(2))) void k_apply(
    KArgs a, const T *pool, const int *seg_list, int seg0, int n, double lo_mult, double hi_mult, int cls_counter,
    unsigned long long *head) {
    __shared__ ApplySmem<N, T> sm;
    extern __shared__ double nrm_s[];
    __shared__ int s_next;
    while (true) {
        const int e = next_entry(head, &s_next);
        if (e >= n) break;
        apply_segment<N, T, PAIR>(a, pool, seg_list ? seg_list[e] : seg0 + e, lo_mult, hi_mult, cls_counter, sm,
                                  nrm_s);
    }
}

// The split class of the split mode: one workgroup per pair segment of
// seg_list (terms with bound <= tau thr_cls, FP16 split products).  It runs
// beside the persistent FP64-class launch on the second stream.  A launch of
// one grid per entry keeps this kernel at 128 VGPRs (4 waves per SIMD at
// N <= 10); a persistent loop around the body, or the FP64 body in the same
// kernel, costs registers and was measured slower.
template <int N>
__global__ __launch_bounds__(SplitShape<N>::T) __attribute__((amdgpu_waves_per_eu(SplitShape<N>::WPE))) void k_apply_h(
    KArgs a, const _Float16 *pool_h, const double *pool_sc, const int *seg_list, double tau, double single_mult) {
    __shared__ SplitSmem<N> sm;
    extern __shared__ double nrm_s[];
    split_segment<N>(a, pool_h, pool_sc, seg_list[blockIdx.x], 0.0, tau, single_mult, 1, sm, nrm_s);
}

// FP16 split rows of FP64 blocks (rows x N values each, row-major), data or
// operator side: one wave per block; scale[b] = 2^e with max |v| < 2^e
// (0 for a zero block).
template <int N, bool DATA>
__global__ void k_split_rows(const double *src, long long nblocks, int rows, _Float16 *dst, double *scale) {
    const long long b = blockIdx.x;
    const int lane = threadIdx.x;
    if (b >= nblocks) return;
    const double *s = src + b * (long long)rows * N;
    double m = 0.0;
    for (int e = lane; e < rows * N; e += 64) m = fmax(m, fabs(s[e]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    int ex = 0;
    if (m > 0.0) (void)frexp(m, &ex);
    const double inv = m > 0.0 ? ldexp(1.0, -ex) : 0.0;
    if (lane == 0) scale[b] = m > 0.0 ? ldexp(1.0, ex) : 0.0;
    for (int r = lane; r < rows; r += 64) {
        double u[12];
#pragma unroll
        for (int jj = 0; jj < 12; jj++) u[jj] = jj < N ? s[r * N + jj] * inv : 0.0;
        split_row<N, DATA>(u, dst + (b * rows + r) * SplitShape<N>::RH);
    }
}

__global__ void k_zero_f64(double *p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// FP64 -> FP32 copy (input pool and operator tables of the FP32 launches)
__global__ void k_to_f32(const double *src, float *dst, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    int i = threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ------------------------------------------------------------------ host side
struct DevOper {
    int *L = nullptr;
    long long *boff = nullptr, *noff = nullptr;
    double *blocks = nullptr, *norms = nullptr;
    float *blocks32 = nullptr;  // FP32 copy, made on first mixed-precision use
    _Float16 *blocks_h = nullptr;  // FP16 split rows, made on first split use
    double *blocks_hsc = nullptr;
    long long nblk = 0;
    int M = 0, ndepth = 0;
};

struct DeviceCtx {
    int device = 0;
    int n_cu = 0;       // compute units
    int lds_max = 0;    // LDS bytes a workgroup may use
    // resident workgroups per CU of a persistent launch, by (kernel, LDS bytes)
    std::map<std::pair<const void *, size_t>, int> occupancy;
    hipStream_t stream = nullptr;
    // second stream (default priority): the FP64-class launch of the split
    // mode runs on it beside the FP16 split launch
    hipStream_t stream2 = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evm = nullptr, evz = nullptr, evl = nullptr, ev2s = nullptr,
               ev2e = nullptr, evp = nullptr;
    // resident operator tables by operator handle; the map is guarded by
    // opers_mtx (an operator's destroy reaches every context from its thread)
    std::map<const void *, DevOper> opers;
    std::mutex opers_mtx;
    // device address and size of the scratch array of the allocation-free
    // entry points (dev_noalloc_init)
    void *noalloc = nullptr;
    size_t noalloc_bytes = 0;
    // pinned staging buffers of the large device-to-host copies (allocated
    // on first use) and the events of their DMAs
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

struct DevicePlan {
    DeviceCtx *ctx = nullptr;
    std::vector<const void *> op_keys;  // per part
    // segments: n_pseg pair ones (indices 0 .. n_pseg - 1), then n_fseg fine
    // single-component ones
    int n_out = 0, kp1 = 0, wmax = 0, n_pseg = 0, n_fseg = 0;
    long long npool = 0;
    double *pool = nullptr;
    float *pool32 = nullptr;
    _Float16 *pool_h = nullptr;  // FP16 split rows of the pool (split mode)
    double *pool_hsc = nullptr;
    KArgs args{};
    std::vector<AppDesc> apps;          // host copy of args.apps
    AppDesc *d_apps = nullptr;
    std::vector<void *> bufs;
    size_t out_bytes = 0;
    // per pair segment: class cut base thr_cls and the range of passing
    // bounds; the same for the fine segments
    std::vector<double> pseg_thr, pseg_bmin, pseg_bmax;
    std::vector<double> fseg_thr, fseg_bmin, fseg_bmax;
    // segment lists of the two classes, for list_tau and list_mode (the FP64
    // class lists fine segments, the low class of the split mode pair ones)
    double list_tau = -1.0;
    int list_mode = -1, list_order = -1;
    // tunable knobs (dev_plan_set_knobs): the split class's single-rounding
    // cut and the order of the work queue (0: segment order, 1: largest
    // passing-term count first)
    double single_ratio = 16.0;
    int order = 0;
    std::vector<int> pseg_apps, fseg_apps;  // passing terms per segment
    int *list_hi = nullptr, *list_lo = nullptr;
    int n_hi = 0, n_lo = 0;
    // post transforms in place (dev_plan_set_post): (branch, first child)
    // pairs by depth, the filter and the node norms
    PostPair *post_pairs = nullptr;
    std::vector<long long> post_off;
    double *post_F = nullptr, *post_norms = nullptr;
    // the captured step (step_graph) and what it was captured for
    hipGraphExec_t graph = nullptr;
    int graph_mode = -1;
    double graph_tau = -1.0;
    bool graph_post = false;
    unsigned ran_eager = 0;  // bit 2 mode + post: an eager run of that step happened
    ~DevicePlan() {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (void *p : bufs)
            if (p) (void)hipFree(p);
        if (pool32) (void)hipFree(pool32);
        if (pool_h) (void)hipFree(pool_h);
        if (pool_hsc) (void)hipFree(pool_hsc);
        if (list_hi) (void)hipFree(list_hi);
        if (list_lo) (void)hipFree(list_lo);
    }
};

int dev_create(int device, DeviceCtx **out, std::string &err) {
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        err = "device index " + std::to_string(device) + " out of range (" + std::to_string(n) + " devices)";
        return -2;
    }
    HIP_TRY(hipSetDevice(device));
    auto *c = new DeviceCtx;
    c->device = device;
    if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        hipDeviceGetAttribute(&c->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) {
        err = "device attribute query failed";
        delete c;
        return -2;
    }
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipEventCreate(&c->evm);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->evz);
    if (e == hipSuccess) e = hipEventCreate(&c->evl);
    if (e == hipSuccess) e = hipEventCreate(&c->ev2s);
    if (e == hipSuccess) e = hipEventCreate(&c->ev2e);
    if (e == hipSuccess) e = hipEventCreate(&c->evp);
    if (e != hipSuccess) {
        err = std::string("stream/event creation: ") + hipGetErrorString(e);
        delete c;
        return -2;
    }
    if (dev_noalloc_init(c, err)) {
        dev_destroy(c);
        return -2;
    }
    *out = c;
    return 0;
}

static void free_oper(DevOper &o) {
    if (o.L) (void)hipFree(o.L);
    if (o.boff) (void)hipFree(o.boff);
    if (o.noff) (void)hipFree(o.noff);
    if (o.blocks) (void)hipFree(o.blocks);
    if (o.norms) (void)hipFree(o.norms);
    if (o.blocks32) (void)hipFree(o.blocks32);
    if (o.blocks_h) (void)hipFree(o.blocks_h);
    if (o.blocks_hsc) (void)hipFree(o.blocks_hsc);
    o = DevOper();
}

void dev_destroy(DeviceCtx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->opers) free_oper(kv.second);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evm) (void)hipEventDestroy(c->evm);
    if (c->evz) (void)hipEventDestroy(c->evz);
    if (c->evl) (void)hipEventDestroy(c->evl);
    if (c->ev2s) (void)hipEventDestroy(c->ev2s);
    if (c->ev2e) (void)hipEventDestroy(c->ev2e);
    if (c->evp) (void)hipEventDestroy(c->evp);
    for (int b = 0; b < 2; b++) {
        if (c->stage[b]) (void)hipHostFree(c->stage[b]);
        if (c->stage_ev[b]) (void)hipEventDestroy(c->stage_ev[b]);
    }
    if (c->stream2) {
        (void)hipStreamSynchronize(c->stream2);
        (void)hipStreamDestroy(c->stream2);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Device-to-host copies into pageable memory run at about 17 GB/s on
// MI355X hosts and DMA into pinned memory at about 57 GB/s, so copies of
// kStageMin bytes or more go through the context's two pinned staging
// buffers: the DMA of chunk i overlaps the host copy of chunk i - 1, which
// kStageThreads threads share.
constexpr size_t kStageChunk = (size_t)64 << 20, kStageMin = (size_t)32 << 20;
constexpr int kStageThreads = 8;

static void par_memcpy(void *dst, const void *src, size_t n) {
    auto part = [&](int t) {
        const size_t a = n * t / kStageThreads, b = n * (t + 1) / kStageThreads;
        std::memcpy(static_cast<char *>(dst) + a, static_cast<const char *>(src) + a, b - a);
    };
    std::thread th[kStageThreads - 1];
    for (int t = 1; t < kStageThreads; t++) th[t - 1] = std::thread(part, t);
    part(0);
    for (auto &x : th) x.join();
}

// The context's two pinned staging buffers and their events, allocated on
// first use; false when pinned memory is refused (the copy then goes
// straight between the device and the pageable host memory).
static bool stage_ready(DeviceCtx *c, std::string &err) {
    for (int b = 0; b < 2; b++) {
        if (!c->stage[b] && hipHostMalloc(&c->stage[b], kStageChunk, hipHostMallocDefault) != hipSuccess) {
            c->stage[b] = nullptr;
            (void)hipGetLastError();
            return false;
        }
    }
    for (int b = 0; b < 2; b++)
        if (!c->stage_ev[b] && hipEventCreateWithFlags(&c->stage_ev[b], hipEventDisableTiming) != hipSuccess) {
            c->stage_ev[b] = nullptr;
            (void)hipGetLastError();
            return false;
        }
    return true;
}

// Host-to-device copies of kStageMin bytes or more go through the same
// staging buffers: the host copy of chunk i into buffer i & 1 overlaps the
// DMA of chunk i - 1 and waits for the DMA of chunk i - 2 (the buffer's
// previous chunk).
int dev_copy_h2d(DeviceCtx *c, void *dst, const void *src, size_t bytes, std::string &err) {
    if (bytes == 0) return 0;
    HIP_TRY(hipSetDevice(c->device));
    if (bytes < kStageMin || !stage_ready(c, err)) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return 0;
    }
    const size_t nch = (bytes + kStageChunk - 1) / kStageChunk;
    auto len = [&](size_t i) { return std::min(kStageChunk, bytes - i * kStageChunk); };
    for (size_t i = 0; i < nch; i++) {
        const int b = (int)(i & 1);
        if (i >= 2) HIP_TRY(hipEventSynchronize(c->stage_ev[b]));
        par_memcpy(c->stage[b], static_cast<const char *>(src) + i * kStageChunk, len(i));
        HIP_TRY(hipMemcpyAsync(static_cast<char *>(dst) + i * kStageChunk, c->stage[b], len(i),
                               hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipEventRecord(c->stage_ev[b], c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dev_copy_d2h(DeviceCtx *c, void *dst, const void *src, size_t bytes, std::string &err) {
    if (bytes == 0) return 0;
    HIP_TRY(hipSetDevice(c->device));
    // without pinned staging memory (allocation refused) the copy goes
    // straight from the device into the pageable destination
    if (bytes < kStageMin || !stage_ready(c, err)) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return 0;
    }
    const size_t nch = (bytes + kStageChunk - 1) / kStageChunk;
    auto len = [&](size_t i) { return std::min(kStageChunk, bytes - i * kStageChunk); };
    // chunk i goes through buffer i & 1; its DMA is queued only after the
    // host has copied chunk i - 2 out of that buffer (the previous iteration)
    for (size_t i = 0; i <= nch; i++) {
        if (i < nch) {
            const int b = (int)(i & 1);
            HIP_TRY(hipMemcpyAsync(c->stage[b], static_cast<const char *>(src) + i * kStageChunk, len(i),
                                   hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipEventRecord(c->stage_ev[b], c->stream));
        }
        if (i > 0) {
            const int b = (int)((i - 1) & 1);
            HIP_TRY(hipEventSynchronize(c->stage_ev[b]));
            par_memcpy(static_cast<char *>(dst) + (i - 1) * kStageChunk, c->stage[b], len(i - 1));
        }
    }
    return 0;
}

void dev_forget_operator(DeviceCtx *c, const void *key) {
    std::lock_guard<std::mutex> g(c->opers_mtx);
    auto it = c->opers.find(key);
    if (it == c->opers.end()) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_oper(it->second);
    c->opers.erase(it);
}

template <typename T>
static int alloc_copy(const std::vector<T> &v, T **out, hipStream_t s, std::string &err) {
    HIP_TRY(hipMalloc((void **)out, sizeof(T) * std::max<size_t>(v.size(), 1)));
    if (!v.empty()) HIP_TRY(hipMemcpyAsync(*out, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
    return 0;
}

static int to_f32(const double *src, float **dst, long long n, hipStream_t s, std::string &err) {
    HIP_TRY(hipMalloc((void **)dst, sizeof(float) * std::max<long long>(n, 1)));
    if (n > 0) {
        const int bs = 256;
        hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, src, *dst, n);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

static int get_oper(DeviceCtx *c, const void *key, const OperFlat &f, DevOper **out, std::string &err) {
    std::lock_guard<std::mutex> g(c->opers_mtx);
    auto it = c->opers.find(key);
    if (it != c->opers.end()) {
        *out = &it->second;
        return 0;
    }
    if (f.device_blocks > 0) {
        err = "operator tables were built on another context's device";
        return -1;
    }
    DevOper o;
    int rc = alloc_copy(f.L, &o.L, c->stream, err);
    if (!rc) rc = alloc_copy(f.boff, &o.boff, c->stream, err);
    if (!rc) rc = alloc_copy(f.noff, &o.noff, c->stream, err);
    if (!rc) rc = alloc_copy(f.blocks, &o.blocks, c->stream, err);
    if (!rc) rc = alloc_copy(f.norms, &o.norms, c->stream, err);
    o.nblk = (long long)f.blocks.size();
    o.M = f.M;
    o.ndepth = f.ndepth;
    if (rc) {
        free_oper(o);
        return rc;
    }
    c->opers[key] = o;
    *out = &c->opers[key];
    return 0;
}

int dev_device(DeviceCtx *c) { return c->device; }
void dev_set_noalloc(DeviceCtx *c, void *base, size_t bytes) {
    c->noalloc = base;
    c->noalloc_bytes = bytes;
}
void *dev_noalloc_base(DeviceCtx *c) { return c->noalloc; }
size_t dev_noalloc_bytes(DeviceCtx *c) { return c->noalloc_bytes; }
void *dev_stream(DeviceCtx *c) { return c->stream; }

int dev_alloc_oper_tables(DeviceCtx *c, const OperFlat &f, DevOperTables &t, std::string &err) {
    t = DevOperTables();
    HIP_TRY(hipMalloc((void **)&t.blocks, sizeof(double) * std::max<long long>(f.device_blocks, 1)));
    HIP_TRY(hipMalloc((void **)&t.norms, sizeof(double) * std::max<size_t>(f.norms.size(), 1)));
    HIP_TRY(hipMemsetAsync(t.blocks, 0, sizeof(double) * std::max<long long>(f.device_blocks, 1), c->stream));
    HIP_TRY(hipMemsetAsync(t.norms, 0, sizeof(double) * std::max<size_t>(f.norms.size(), 1), c->stream));
    int rc = alloc_copy(f.L, &t.L, c->stream, err);
    if (!rc) rc = alloc_copy(f.boff, &t.boff, c->stream, err);
    if (!rc) rc = alloc_copy(f.noff, &t.noff, c->stream, err);
    return rc;
}

int dev_register_operator(DeviceCtx *c, const void *key, const OperFlat &f, const DevOperTables &t, std::string &err) {
    (void)err;
    std::lock_guard<std::mutex> g(c->opers_mtx);
    auto it = c->opers.find(key);
    if (it != c->opers.end()) {
        (void)hipStreamSynchronize(c->stream);
        free_oper(it->second);
        c->opers.erase(it);
    }
    DevOper o;
    o.L = t.L;
    o.boff = t.boff;
    o.noff = t.noff;
    o.blocks = t.blocks;
    o.norms = t.norms;
    o.nblk = f.device_blocks;
    o.M = f.M;
    o.ndepth = f.ndepth;
    c->opers[key] = o;
    return 0;
}

int dev_download_operator_blocks(DeviceCtx *c, const void *key, double *out, long long n, std::string &err) {
    HIP_TRY(hipSetDevice(c->device));
    std::lock_guard<std::mutex> g(c->opers_mtx);
    auto it = c->opers.find(key);
    if (it == c->opers.end()) {
        err = "operator tables are not resident on this context";
        return -1;
    }
    if (n != it->second.nblk) {
        err = "operator table size mismatch";
        return -1;
    }
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(out, it->second.blocks, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return 0;
}

// Device buffer of n elements owned by the plan.
template <typename T>
static int plan_alloc(DevicePlan *p, size_t n, T **out, std::string &err) {
    void *d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(T) * std::max<size_t>(n, 1)));
    p->bufs.push_back(d);
    *out = static_cast<T *>(d);
    return 0;
}

template <typename T, typename A>
static int upload_at(T *dst, const std::vector<T, A> &v, hipStream_t s, std::string &err) {
    if (!v.empty()) HIP_TRY(hipMemcpyAsync(dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
    return 0;
}

int dev_plan_create(DeviceCtx *c, const std::vector<PlanPart> &parts, DevicePlan **out, std::string &err,
                    double *adopt_pool) {
    HIP_TRY(hipSetDevice(c->device));
    if (parts.empty()) {
        err = "empty plan";
        return -1;
    }
    const int N = parts[0].plan->kp1;
    if (N < 2 || N > 12) {
        err = "unsupported basis size k+1 = " + std::to_string(N);
        return -1;
    }
    if ((long long)(2 * kMaxBand + 1) * (2 * kMaxBand + 1) * (2 * kMaxBand + 1) >= (1 << 22)) {
        err = "band index does not fit the worklist encoding";
        return -1;
    }
    // sizes and validation
    long long n_out = 0, n_band = 0, n_slot = 0, n_pool = 0, n_bw = 0, n_actoff = 0, n_act = 0;
    long long n_pseg = 0, n_fseg = 0;
    int wmax = 0;
    for (const PlanPart &pp : parts) {
        const ApplyPlan &pl = *pp.plan;
        if (pl.kp1 != N) {
            err = "all applies of a batch must share one basis size";
            return -1;
        }
        for (size_t d = 0; d + 1 < pl.active_off.size(); d++)
            if (pl.active_off[d + 1] - pl.active_off[d] > kMaxActive) {
                err = "more than " + std::to_string(kMaxActive) + " active terms at one depth";
                return -1;
            }
        for (int wq : pl.band_w) {
            if (wq > kMaxBand) {
                err = "band half-width " + std::to_string(wq) + " exceeds the kernel limit " + std::to_string(kMaxBand);
                return -1;
            }
            wmax = std::max(wmax, wq);
        }
        n_out += (long long)pl.out_node.size();
        n_band += (long long)pl.band.size();
        n_slot += (long long)pl.slot_off.size();
        n_pool += pl.device_pool ? pl.pool_values : (long long)pl.in_coefs.size();
        n_bw += (long long)pl.bw.size();
        n_actoff += (long long)pl.active_off.size();
        n_act += (long long)pl.active_terms.size();
        n_pseg += (long long)pl.pseg_job.size();
        n_fseg += (long long)pl.fseg_job.size();
    }
    const long long n_all = n_pseg + n_fseg;
    if (n_out >= (1LL << 27) || n_all >= (1LL << 31) || n_slot >= (1LL << 31) || n_band >= (1LL << 31)) {
        err = "batch too large for 32-bit node, segment or slot indices";
        return -1;
    }
    auto *p = new DevicePlan;
    p->ctx = c;
    p->n_out = (int)n_out;
    p->kp1 = N;
    p->wmax = wmax;
    p->npool = n_pool;
    p->n_pseg = (int)n_pseg;
    p->n_fseg = (int)n_fseg;
    KArgs &a = p->args;
    hipStream_t s = c->stream;
    int rc = 0;
    // Every table of the plan but the pool and the output lives in one
    // device arena, filled from one host staging buffer with one copy: a
    // refinement level's plan costs one allocation and one upload however
    // many applies it holds.
    std::vector<DevOper *> opers(parts.size(), nullptr);
    std::vector<std::vector<double>> mtabs;
    std::vector<std::pair<int, int>> mt_shape;
    for (size_t ai = 0; ai < parts.size() && !rc; ai++) rc = get_oper(c, parts[ai].op_key, *parts[ai].flat, &opers[ai], err);
    if (rc) {
        delete p;
        return rc;
    }
    {
        std::vector<const ApplyPlan *> pls;
        std::vector<const OperFlat *> fs;
        for (const PlanPart &pp : parts) {
            pls.push_back(pp.plan);
            fs.push_back(pp.flat);
        }
        build_prescreens(pls, fs, mtabs, mt_shape);
    }
    size_t arena = 0;
    auto place = [&](size_t bytes) {
        const size_t off = arena;
        arena += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        return off;
    };
    const size_t f_out_depth = place(sizeof(int) * n_out), f_out_app = place(sizeof(int) * n_out);
    const size_t f_thr = place(sizeof(double) * n_out), f_band_off = place(sizeof(long long) * n_out);
    const size_t f_band_w = place(sizeof(int) * n_out), f_band = place(sizeof(int) * n_band);
    const size_t f_slot_off = place(sizeof(long long) * n_slot), f_norm = place(sizeof(double) * 8 * n_slot);
    const size_t f_bw = place(sizeof(int) * n_bw), f_act_off = place(sizeof(int) * n_actoff);
    const size_t f_act = place(sizeof(int) * n_act);
    const size_t f_sj = place(sizeof(int) * n_all), f_s0 = place(sizeof(int) * n_all), f_s1 = place(sizeof(int) * n_all);
    const size_t f_apps = place(sizeof(AppDesc) * parts.size());
    const size_t f_cnt = place(sizeof(unsigned long long) * 8);
    std::vector<size_t> f_mt(parts.size());
    for (size_t ai = 0; ai < parts.size(); ai++) f_mt[ai] = place(sizeof(double) * mtabs[ai].size());
    if (adopt_pool) {
        // the parts' pools must sit one after another in it
        long long off = 0;
        for (const PlanPart &pp : parts) {
            if (!pp.plan->device_pool || pp.dev_pool != adopt_pool + off) {
                err = "adopted pool does not hold the parts' pools in order";
                delete p;
                return -1;
            }
            off += pp.plan->pool_values;
        }
    }
    unsigned char *dev = nullptr;
    rc = plan_alloc(p, arena, &dev, err);
    if (!rc && adopt_pool) {
        p->pool = adopt_pool;  // owned (in bufs) only once the plan is made
    } else if (!rc) {
        rc = plan_alloc(p, (size_t)std::max<long long>(n_pool, 1), &p->pool, err);
    }
    p->out_bytes = sizeof(double) * (size_t)std::max<long long>(n_out, 1) * 8 * N * N * N;
    if (!rc) rc = plan_alloc(p, p->out_bytes / sizeof(double), &a.out_ns, err);
    if (rc) {
        delete p;
        return rc;
    }
    std::vector<unsigned char, UninitAlloc<unsigned char>> stage(arena);
    auto at = [&](size_t f) { return stage.data() + f; };
    auto put = [&](size_t f, long long off, const auto &v) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        if (!v.empty()) std::memcpy(at(f) + sizeof(T) * off, v.data(), sizeof(T) * v.size());
    };
    long long o_out = 0, o_band = 0, o_slot = 0, o_pool = 0, o_bw = 0, o_actoff = 0, o_act = 0;
    long long o_pseg = 0, o_fseg = n_pseg;
    // a segment list's jobs, shifted to the batch's output numbering
    auto put_jobs = [&](long long o_dst, const std::vector<int> &jobs, const std::vector<int> &p0,
                        const std::vector<int> &p1, long long o_o) {
        int *sj = reinterpret_cast<int *>(at(f_sj)) + o_dst;
        for (size_t i = 0; i < jobs.size(); i++) sj[i] = jobs[i] + (int)(8 * o_o);
        put(f_s0, o_dst, p0);
        put(f_s1, o_dst, p1);
    };
    for (size_t ai = 0; ai < parts.size() && !rc; ai++) {
        const PlanPart &pp = parts[ai];
        const ApplyPlan &pl = *pp.plan;
        DevOper *o = opers[ai];
        p->op_keys.push_back(pp.op_key);
        AppDesc ad{};
        ad.opL = o->L;
        ad.op_boff = o->boff;
        ad.op_noff = o->noff;
        ad.op_norms = o->norms;
        ad.blk64 = o->blocks;
        ad.blk32 = o->blocks32;
        ad.blk_h = o->blocks_h;
        ad.blk_hsc = o->blocks_hsc;
        ad.M = pp.flat->M;
        ad.ndepth = pp.flat->ndepth;
        ad.bw_off = o_bw;
        ad.act_base = (int)o_actoff;
        ad.thr_floor = pl.thr_floor;
        ad.mt_d0 = mt_shape[ai].first;
        ad.mt_w = mt_shape[ai].second;
        put(f_mt[ai], 0, mtabs[ai]);
        ad.mtab = reinterpret_cast<const double *>(dev + f_mt[ai]);
        p->apps.push_back(ad);
        const size_t no = pl.out_node.size();
        // outputs
        put(f_out_depth, o_out, pl.out_depth);
        std::fill_n(reinterpret_cast<int *>(at(f_out_app)) + o_out, no, (int)ai);
        put(f_thr, o_out, pl.out_thr);
        {
            long long *bo = reinterpret_cast<long long *>(at(f_band_off)) + o_out;
            for (size_t i = 0; i < no; i++) bo[i] = pl.band_off[i] + o_band;
            int *bd = reinterpret_cast<int *>(at(f_band)) + o_band;
            for (size_t i = 0; i < pl.band.size(); i++) bd[i] = pl.band[i] >= 0 ? pl.band[i] + (int)o_slot : pl.band[i];
        }
        put(f_band_w, o_out, pl.band_w);
        // inputs
        {
            long long *so = reinterpret_cast<long long *>(at(f_slot_off)) + o_slot;
            for (size_t i = 0; i < pl.slot_off.size(); i++) so[i] = pl.slot_off[i] + o_pool;
        }
        if (pl.device_pool) {
            if (!pp.dev_pool) {
                err = "plan part without its device pool";
                rc = -1;
            } else if (pl.pool_values > 0 && !adopt_pool) {
                hipError_t e = hipMemcpyAsync(p->pool + o_pool, pp.dev_pool, sizeof(double) * pl.pool_values,
                                              hipMemcpyDeviceToDevice, s);
                if (e != hipSuccess) {
                    err = std::string("pool copy: ") + hipGetErrorString(e);
                    rc = -2;
                }
            }
        } else {
            rc = upload_at(p->pool + o_pool, pl.in_coefs, s, err);
        }
        put(f_norm, o_slot * 8, pl.in_norms);
        // screening tables
        put(f_bw, o_bw, pl.bw);
        {
            int *ao = reinterpret_cast<int *>(at(f_act_off)) + o_actoff;
            for (size_t i = 0; i < pl.active_off.size(); i++) ao[i] = pl.active_off[i] + (int)o_act;
        }
        put(f_act, o_act, pl.active_terms);
        // segments
        p->pseg_apps.insert(p->pseg_apps.end(), pl.pseg_apps.begin(), pl.pseg_apps.end());
        p->fseg_apps.insert(p->fseg_apps.end(), pl.fseg_apps.begin(), pl.fseg_apps.end());
        put_jobs(o_pseg, pl.pseg_job, pl.pseg_p0, pl.pseg_p1, o_out);
        for (size_t i = 0; i < pl.pseg_job.size(); i++) {
            const double t = pl.out_thr[pl.pseg_job[i] >> 3];
            p->pseg_thr.push_back(t > pl.thr_floor ? t : pl.thr_floor);  // as the kernels form thr_cls
            p->pseg_bmin.push_back(pl.pseg_bmin[i]);
            p->pseg_bmax.push_back(pl.pseg_bmax[i]);
        }
        o_pseg += (long long)pl.pseg_job.size();
        put_jobs(o_fseg, pl.fseg_job, pl.fseg_p0, pl.fseg_p1, o_out);
        for (size_t i = 0; i < pl.fseg_job.size(); i++) {
            const double t = pl.out_thr[pl.fseg_job[i] >> 3];
            p->fseg_thr.push_back(t > pl.thr_floor ? t : pl.thr_floor);
            p->fseg_bmin.push_back(pl.fseg_bmin[i]);
            p->fseg_bmax.push_back(pl.fseg_bmax[i]);
        }
        o_fseg += (long long)pl.fseg_job.size();
        o_out += (long long)no;
        o_band += (long long)pl.band.size();
        o_slot += (long long)pl.slot_off.size();
        o_pool += pl.device_pool ? pl.pool_values : (long long)pl.in_coefs.size();
        o_bw += (long long)pl.bw.size();
        o_actoff += (long long)pl.active_off.size();
        o_act += (long long)pl.active_terms.size();
    }
    if (!rc) {
        std::memcpy(at(f_apps), p->apps.data(), sizeof(AppDesc) * p->apps.size());
        std::memset(at(f_cnt), 0, sizeof(unsigned long long) * 8);
        hipError_t e = hipMemcpyAsync(dev, stage.data(), arena, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            err = std::string("upload: ") + hipGetErrorString(e);
            rc = -2;
        }
    }
    if (rc) {
        (void)hipStreamSynchronize(s);
        delete p;
        return rc;
    }
    p->d_apps = reinterpret_cast<AppDesc *>(dev + f_apps);
    a.counters = reinterpret_cast<unsigned long long *>(dev + f_cnt);
    a.out_depth = reinterpret_cast<const int *>(dev + f_out_depth);
    a.out_app = reinterpret_cast<const int *>(dev + f_out_app);
    a.out_thr = reinterpret_cast<const double *>(dev + f_thr);
    a.band_off = reinterpret_cast<const long long *>(dev + f_band_off);
    a.band_w = reinterpret_cast<const int *>(dev + f_band_w);
    a.band = reinterpret_cast<const int *>(dev + f_band);
    a.slot_off = reinterpret_cast<const long long *>(dev + f_slot_off);
    a.slot_norm = reinterpret_cast<const double *>(dev + f_norm);
    a.bw = reinterpret_cast<const int *>(dev + f_bw);
    a.act_off = reinterpret_cast<const int *>(dev + f_act_off);
    a.act_terms = reinterpret_cast<const int *>(dev + f_act);
    a.apps = p->d_apps;
    a.seg_job = reinterpret_cast<const int *>(dev + f_sj);
    a.seg_p0 = reinterpret_cast<const int *>(dev + f_s0);
    a.seg_p1 = reinterpret_cast<const int *>(dev + f_s1);
    if (adopt_pool) p->bufs.push_back(adopt_pool);
    *out = p;
    return 0;
}

int dev_plan_set_post(DevicePlan *p, const MRA &mra,
                      const std::vector<std::vector<std::pair<long long, long long>>> &levels, std::string &err) {
    DeviceCtx *c = p->ctx;
    HIP_TRY(hipSetDevice(c->device));
    if (p->post_pairs) {
        err = "post transforms already set";
        return -1;
    }
    const int N = p->kp1;
    const size_t blk = (size_t)8 * N * N * N;
    std::vector<PostPair> pairs;
    p->post_off.assign(1, 0);
    for (const auto &l : levels) {
        for (const auto &pr : l) {
            if (pr.first < 0 || pr.first >= p->n_out || pr.second < 0 || pr.second + 8 > p->n_out) {
                err = "post transform pair outside the plan's output nodes";
                return -1;
            }
            pairs.push_back(PostPair{p->args.out_ns + (size_t)pr.first * blk, p->args.out_ns + (size_t)pr.second * blk});
        }
        p->post_off.push_back((long long)pairs.size());
    }
    int rc = plan_alloc(p, pairs.size(), &p->post_pairs, err);
    if (!rc) rc = plan_alloc(p, mra.basis.F.size(), &p->post_F, err);
    if (!rc) rc = plan_alloc(p, 2 * (size_t)std::max(p->n_out, 1), &p->post_norms, err);
    if (!rc) rc = upload_at(p->post_pairs, pairs, c->stream, err);
    if (!rc) rc = upload_at(p->post_F, mra.basis.F, c->stream, err);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dev_plan_download_norms(DevicePlan *p, double *s2, double *w2, std::string &err) {
    DeviceCtx *c = p->ctx;
    HIP_TRY(hipSetDevice(c->device));
    if (!p->post_norms) {
        err = "the plan has no post transforms";
        return -1;
    }
    if (p->n_out == 0) return 0;
    HIP_TRY(hipMemcpyAsync(s2, p->post_norms, sizeof(double) * p->n_out, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(w2, p->post_norms + p->n_out, sizeof(double) * p->n_out, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

void dev_plan_destroy(DevicePlan *p) {
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    delete p;
}

// Norm table bytes at the start of a job's LDS (16-byte aligned end).
static int nrm_bytes(int wmax) { return (int)((sizeof(double) * 64 * nrm_stride(2 * wmax + 1) + 15) / 16 * 16); }

// Grid of a persistent launch: the resident workgroups of `kernel` with
// `static_lds` bytes of static and `shmem` bytes of dynamic LDS (occupancy
// per CU x CUs), at most n.
static int persistent_grid(DeviceCtx *c, const void *kernel, int threads, size_t static_lds, size_t shmem, int n,
                           int *grid, std::string &err) {
    if (static_lds + shmem + 16 > (size_t)c->lds_max) {
        err = "a job needs " + std::to_string(static_lds + shmem) + " bytes of LDS, more than the " +
              std::to_string(c->lds_max) + " a workgroup may use (band too wide for this basis size)";
        return -1;
    }
    auto key = std::make_pair(kernel, shmem);
    auto it = c->occupancy.find(key);
    int per_cu = 0;
    if (it == c->occupancy.end()) {
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, shmem));
        c->occupancy[key] = per_cu;
    } else {
        per_cu = it->second;
    }
    if (per_cu < 1) {
        err = "the apply kernel does not fit on a compute unit";
        return -1;
    }
    *grid = std::min(n, per_cu * c->n_cu);
    return 0;
}

template <int N, typename T, bool PAIR>
static int launch_nt(DeviceCtx *c, const KArgs &a, const T *pool, const int *list, int seg0, int n, int wmax,
                     double lo, double hi, int cls, unsigned long long *head, hipStream_t s, std::string &err) {
    if (n == 0) return 0;
    const size_t shmem = nrm_bytes(wmax);
    int grid = 0;
    int rc = persistent_grid(c, reinterpret_cast<const void *>(&k_apply<N, T, PAIR>), JobShape<N>::T,
                             sizeof(ApplySmem<N, T>), shmem, n, &grid, err);
    if (rc) return rc;
    hipLaunchKernelGGL((k_apply<N, T, PAIR>), dim3(grid), dim3(JobShape<N>::T), shmem, s, a, pool, list, seg0, n, lo,
                       hi, cls, head);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Persistent launch of k_apply over seg_list[0, n) or, without a list,
// segments seg0 .. seg0 + n - 1; pair: the segments are pair segments.
template <typename T>
static int launch_t(const DevicePlan *p, const T *pool, bool pair, const int *list, int seg0, int n, double lo,
                    double hi, int cls, unsigned long long *head, std::string &err, hipStream_t s = nullptr) {
    const KArgs &a = p->args;
    if (!s) s = p->ctx->stream;
#define MWG_LAUNCH(NN)                                                                                           \
    (pair ? launch_nt<NN, T, true>(p->ctx, a, pool, list, seg0, n, p->wmax, lo, hi, cls, head, s, err)           \
          : launch_nt<NN, T, false>(p->ctx, a, pool, list, seg0, n, p->wmax, lo, hi, cls, head, s, err))
    switch (p->kp1) {
        case 2: return MWG_LAUNCH(2);
        case 3: return MWG_LAUNCH(3);
        case 4: return MWG_LAUNCH(4);
        case 5: return MWG_LAUNCH(5);
        case 6: return MWG_LAUNCH(6);
        case 7: return MWG_LAUNCH(7);
        case 8: return MWG_LAUNCH(8);
        case 9: return MWG_LAUNCH(9);
        case 10: return MWG_LAUNCH(10);
        case 11: return MWG_LAUNCH(11);
        case 12: return MWG_LAUNCH(12);
        default: err = "unsupported basis size"; return -1;
    }
#undef MWG_LAUNCH
}

// Segment lists of the mixed mode: segments holding a term with bound above
// tau * thr_cls (the FP64 class) and segments holding one at or below it.  The
// lists are widened by a relative 1e-9 so that rounding can only add segments;
// the kernel's own class test decides each term.
static int drop_graph(DevicePlan *p, std::string &err);

static int make_lists(DevicePlan *p, int mode, double tau, std::string &err) {
    if (p->list_tau == tau && p->list_mode == mode && p->list_order == p->order) return 0;
    std::vector<int> hi, lo;
    for (int i = 0; i < p->n_fseg; i++)
        if (p->fseg_bmax[i] > tau * p->fseg_thr[i] * (1.0 - 1e-9)) hi.push_back(p->n_pseg + i);
    if (mode == 1 || mode == 2)
        for (int i = 0; i < p->n_pseg; i++)
            if (p->pseg_bmin[i] <= tau * p->pseg_thr[i] * (1.0 + 1e-9)) lo.push_back(i);
    if (p->order == 1) {
        // largest first (segment index -> its passing-term count)
        auto apps = [&](int sg) {
            if (sg < p->n_pseg) return p->pseg_apps[sg];
            return p->fseg_apps[sg - p->n_pseg];
        };
        auto by_cost = [&](int x, int y) { return apps(x) > apps(y); };
        std::stable_sort(hi.begin(), hi.end(), by_cost);
        std::stable_sort(lo.begin(), lo.end(), by_cost);
    }
    // a captured step holds the old lists
    int rc0 = drop_graph(p, err);
    if (rc0) return rc0;
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    if (p->list_hi) HIP_TRY(hipFree(p->list_hi));
    if (p->list_lo) HIP_TRY(hipFree(p->list_lo));
    p->list_hi = p->list_lo = nullptr;
    p->list_tau = -1.0;
    int rc = alloc_copy(hi, &p->list_hi, p->ctx->stream, err);
    if (!rc) rc = alloc_copy(lo, &p->list_lo, p->ctx->stream, err);
    if (rc) return rc;
    p->n_hi = (int)hi.size();
    p->n_lo = (int)lo.size();
    p->list_tau = tau;
    p->list_mode = mode;
    p->list_order = p->order;
    return 0;
}

// FP32 copies of the pool and of every part's operator blocks.
static int ensure_f32(DevicePlan *p, std::string &err) {
    DeviceCtx *c = p->ctx;
    bool changed = false;
    std::lock_guard<std::mutex> g(c->opers_mtx);
    for (size_t i = 0; i < p->op_keys.size(); i++) {
        auto it = c->opers.find(p->op_keys[i]);
        if (it == c->opers.end()) {
            err = "operator tables not resident (operator destroyed while a plan uses it)";
            return -1;
        }
        DevOper &o = it->second;
        if (!o.blocks32) {
            int rc = to_f32(o.blocks, &o.blocks32, o.nblk, c->stream, err);
            if (rc) return rc;
        }
        if (p->apps[i].blk32 != o.blocks32) {
            p->apps[i].blk32 = o.blocks32;
            changed = true;
        }
    }
    if (changed) {
        int rc = upload_at(p->d_apps, p->apps, c->stream, err);
        if (rc) return rc;
    }
    if (!p->pool32) return to_f32(p->pool, &p->pool32, p->npool, c->stream, err);
    return 0;
}

// halfs per split row for basis size N
static int split_rh(int N) { return N > 10 ? 32 : 24; }

template <int N>
static int split_rows_n(const double *src, long long nblocks, int rows, bool data, _Float16 **dst, double **scale,
                        hipStream_t s, std::string &err) {
    HIP_TRY(hipMalloc((void **)dst, sizeof(_Float16) * (size_t)std::max<long long>(nblocks, 1) * rows * SplitShape<N>::RH));
    HIP_TRY(hipMalloc((void **)scale, sizeof(double) * (size_t)std::max<long long>(nblocks, 1)));
    if (nblocks == 0) return 0;
    if (data)
        hipLaunchKernelGGL((k_split_rows<N, true>), dim3((unsigned)nblocks), dim3(64), 0, s, src, nblocks, rows, *dst,
                           *scale);
    else
        hipLaunchKernelGGL((k_split_rows<N, false>), dim3((unsigned)nblocks), dim3(64), 0, s, src, nblocks, rows, *dst,
                           *scale);
    HIP_TRY(hipGetLastError());
    return 0;
}

// FP16 split rows of FP64 blocks of `rows` x N values (data or operator side).
static int split_rows(int N, const double *src, long long nblocks, int rows, bool data, _Float16 **dst,
                      double **scale, hipStream_t s, std::string &err) {
    if (nblocks + 1 >= (1LL << 31)) {
        err = "too many blocks for the FP16 split conversion";
        return -1;
    }
    switch (N) {
        case 2: return split_rows_n<2>(src, nblocks, rows, data, dst, scale, s, err);
        case 3: return split_rows_n<3>(src, nblocks, rows, data, dst, scale, s, err);
        case 4: return split_rows_n<4>(src, nblocks, rows, data, dst, scale, s, err);
        case 5: return split_rows_n<5>(src, nblocks, rows, data, dst, scale, s, err);
        case 6: return split_rows_n<6>(src, nblocks, rows, data, dst, scale, s, err);
        case 7: return split_rows_n<7>(src, nblocks, rows, data, dst, scale, s, err);
        case 8: return split_rows_n<8>(src, nblocks, rows, data, dst, scale, s, err);
        case 9: return split_rows_n<9>(src, nblocks, rows, data, dst, scale, s, err);
        case 10: return split_rows_n<10>(src, nblocks, rows, data, dst, scale, s, err);
        case 11: return split_rows_n<11>(src, nblocks, rows, data, dst, scale, s, err);
        case 12: return split_rows_n<12>(src, nblocks, rows, data, dst, scale, s, err);
        default: err = "unsupported basis size"; return -1;
    }
}

// Split-mode items whose bound is at most single_ratio x thr_cls (default
// kSingleRatio; a tunable knob of the plan) keep their
// intermediates rounded once to FP16 (relative error ~1e-3 of a term bounded
// by 16 thresholds); larger ones use hi/lo intermediates.
constexpr double kSingleRatio = 16.0;

// FP16 split copies of the pool and of every part's operator blocks.
static int ensure_h16(DevicePlan *p, std::string &err) {
    DeviceCtx *c = p->ctx;
    const int N = p->kp1;
    bool changed = false;
    std::lock_guard<std::mutex> g(c->opers_mtx);
    for (size_t i = 0; i < p->op_keys.size(); i++) {
        auto it = c->opers.find(p->op_keys[i]);
        if (it == c->opers.end()) {
            err = "operator tables not resident (operator destroyed while a plan uses it)";
            return -1;
        }
        DevOper &o = it->second;
        if (!o.blocks_h) {
            const long long nb = o.nblk / ((long long)N * N);
            int rc = split_rows(N, o.blocks, nb, N, false, &o.blocks_h, &o.blocks_hsc, c->stream, err);
            if (rc) return rc;
        }
        if (p->apps[i].blk_h != o.blocks_h) {
            p->apps[i].blk_h = o.blocks_h;
            p->apps[i].blk_hsc = o.blocks_hsc;
            changed = true;
        }
    }
    if (changed) {
        int rc = upload_at(p->d_apps, p->apps, c->stream, err);
        if (rc) return rc;
    }
    if (!p->pool_h) {
        const long long nb = p->npool / ((long long)N * N * N);
        return split_rows(N, p->pool, nb, N * N, true, &p->pool_h, &p->pool_hsc, c->stream, err);
    }
    return 0;
}

template <int N>
static int launch_h_n(const DevicePlan *p, double tau, std::string &err) {
    if (p->n_lo == 0) return 0;
    const size_t shmem = nrm_bytes(p->wmax);
    if (sizeof(SplitSmem<N>) + shmem + 16 > (size_t)p->ctx->lds_max) {
        err = "band half-width " + std::to_string(p->wmax) + " needs more LDS than the split kernel has at k+1 = " +
              std::to_string(N) + "; use the FP64 or mixed mode";
        return -1;
    }
    hipLaunchKernelGGL((k_apply_h<N>), dim3(p->n_lo), dim3(SplitShape<N>::T), shmem, p->ctx->stream, p->args,
                       p->pool_h, p->pool_hsc, p->list_lo, tau, p->single_ratio);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int launch_h(const DevicePlan *p, double tau, std::string &err) {
    switch (p->kp1) {
        case 2: return launch_h_n<2>(p, tau, err);
        case 3: return launch_h_n<3>(p, tau, err);
        case 4: return launch_h_n<4>(p, tau, err);
        case 5: return launch_h_n<5>(p, tau, err);
        case 6: return launch_h_n<6>(p, tau, err);
        case 7: return launch_h_n<7>(p, tau, err);
        case 8: return launch_h_n<8>(p, tau, err);
        case 9: return launch_h_n<9>(p, tau, err);
        case 10: return launch_h_n<10>(p, tau, err);
        case 11: return launch_h_n<11>(p, tau, err);
        case 12: return launch_h_n<12>(p, tau, err);
        default: err = "unsupported basis size"; return -1;
    }
}

// The launch sequence of one step on the context's stream(s): counters and
// output cleared by kernels (no memset nodes, so the sequence can be
// captured into a graph), the class launches of the mode, and with post the
// TopDown/BottomUp transforms and the node norms.  With timed the events
// of the per-class and post times are recorded between the launches (not
// during a capture, where only the fork/join events of the second stream
// are recorded).
static int enqueue_step(DevicePlan *p, int mode, double tau, bool post, bool timed, std::string &err) {
    DeviceCtx *c = p->ctx;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, p->args.counters, 8);
    HIP_TRY(hipGetLastError());
    {
        const long long nz = (long long)p->n_out * 8 * p->kp1 * p->kp1 * p->kp1;
        const int bs = 256;
        hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((nz + bs - 1) / bs)), dim3(bs), 0, c->stream, p->args.out_ns, nz);
        HIP_TRY(hipGetLastError());
    }
    int rc = 0;
    const double inf = HUGE_VAL;
    // queue heads of the persistent launches: counters[2] (first stream),
    // counters[3] (second stream)
    unsigned long long *head1 = p->args.counters + 2, *head2 = p->args.counters + 3;
    if (mode == 0) {
        // lo = 0: every term that passes screening, over the pair segments
        rc = launch_t<double>(p, p->pool, true, nullptr, 0, p->n_pseg, 0.0, inf, 0, head1, err);
        if (rc) return rc;
    } else if (mode == 1) {
        // The FP64 class, then the FP32 class, on the first stream; both only
        // add into the zeroed output.  The two persistent launches one after
        // the other take less wall time than the FP64 class on the second
        // stream beside the FP32 class: run side by side, the two grids share
        // the CUs and the FP64 class (a few percent of the term-apps) ends
        // with the FP32 class instead of before it (64 Helmholtz applies at
        // k = 11, 119,968 output nodes: 316 ms against 394 ms a step).
        if (timed) HIP_TRY(hipEventRecord(c->ev2s, c->stream));
        rc = launch_t<double>(p, p->pool, false, p->list_hi, 0, p->n_hi, tau, inf, 0, head2, err, c->stream);
        if (rc) return rc;
        if (timed) HIP_TRY(hipEventRecord(c->ev2e, c->stream));
        if (timed) HIP_TRY(hipEventRecord(c->evl, c->stream));
        rc = launch_t<float>(p, p->pool32, true, p->list_lo, 0, p->n_lo, 0.0, tau, 1, head1, err);
        if (rc) return rc;
        if (timed) HIP_TRY(hipEventRecord(c->evm, c->stream));
    } else {
        // the FP64 class (persistent, fine segments) on the second stream
        // beside the split class on the first, as in the mixed mode
        HIP_TRY(hipEventRecord(c->evz, c->stream));
        HIP_TRY(hipStreamWaitEvent(c->stream2, c->evz, 0));
        if (timed) HIP_TRY(hipEventRecord(c->ev2s, c->stream2));
        rc = launch_t<double>(p, p->pool, false, p->list_hi, 0, p->n_hi, tau, inf, 0, head2, err, c->stream2);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(c->ev2e, c->stream2));
        if (timed) HIP_TRY(hipEventRecord(c->evl, c->stream));
        rc = launch_h(p, tau, err);
        if (rc) return rc;
        if (timed) HIP_TRY(hipEventRecord(c->evm, c->stream));
        HIP_TRY(hipStreamWaitEvent(c->stream, c->ev2e, 0));
    }
    // end of the calculator (both classes joined)
    if (timed) HIP_TRY(hipEventRecord(c->evp, c->stream));
    if (post) {
        // TopDown/BottomUp in place on the output blocks, then the node norms
        rc = dev_post_launch(c->stream, p->kp1, p->post_pairs, p->post_off, p->post_F, err);
        if (!rc)
            rc = dev_node_norms_launch(c->stream, p->kp1, p->args.out_ns, p->n_out, p->post_norms,
                                       p->post_norms + p->n_out, err);
        if (rc) return rc;
    }
    return 0;
}

static int drop_graph(DevicePlan *p, std::string &err) {
    if (p->graph) {
        HIP_TRY(hipGraphExecDestroy(p->graph));
        p->graph = nullptr;
        p->graph_mode = -1;
    }
    return 0;
}

int dev_plan_set_knobs(DevicePlan *p, double single_ratio, int order, std::string &err) {
    if (!(single_ratio >= 0.0) || !std::isfinite(single_ratio) || (order != 0 && order != 1)) {
        err = "knobs: single-rounding ratio must be finite and >= 0, order 0 or 1";
        return -1;
    }
    if (single_ratio != p->single_ratio || order != p->order) {
        HIP_TRY(hipSetDevice(p->ctx->device));
        int rc = drop_graph(p, err);
        if (rc) return rc;
        p->single_ratio = single_ratio;
        p->order = order;
    }
    return 0;
}

void dev_plan_knobs(const DevicePlan *p, double *single_ratio, int *order) {
    *single_ratio = p->single_ratio;
    *order = p->order;
}

int dev_plan_snapshot(DevicePlan *p, double **copy, std::string &err) { return dev_plan_clone_output(p, copy, err); }

// Largest |a - b| and largest |b| over n doubles: per-workgroup maxima
// (grid-stride), then one workgroup over those.
__global__ __launch_bounds__(256) void k_max_diff(const double *a, const double *b, long long n, double *part) {
    __shared__ double sd[256], sb[256];
    double md = 0.0, mb = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        md = fmax(md, fabs(a[i] - b[i]));
        mb = fmax(mb, fabs(b[i]));
    }
    sd[threadIdx.x] = md;
    sb[threadIdx.x] = mb;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + o]);
            sb[threadIdx.x] = fmax(sb[threadIdx.x], sb[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sd[0];
        part[2 * blockIdx.x + 1] = sb[0];
    }
}

int dev_plan_max_diff(DevicePlan *p, const double *ref, double *max_diff, double *max_ref, std::string &err) {
    DeviceCtx *c = p->ctx;
    HIP_TRY(hipSetDevice(c->device));
    const long long n = (long long)p->n_out * 8 * p->kp1 * p->kp1 * p->kp1;
    constexpr int nb = 1024;
    double *part = nullptr;
    HIP_TRY(hipMalloc((void **)&part, sizeof(double) * 2 * nb));
    hipLaunchKernelGGL(k_max_diff, dim3(nb), dim3(256), 0, c->stream, p->args.out_ns, ref, n, part);
    std::vector<double> h(2 * nb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(part);
    if (e != hipSuccess) {
        err = std::string("max difference: ") + hipGetErrorString(e);
        return -2;
    }
    *max_diff = *max_ref = 0.0;
    for (int i = 0; i < nb; i++) {
        *max_diff = std::max(*max_diff, h[2 * i]);
        *max_ref = std::max(*max_ref, h[2 * i + 1]);
    }
    return 0;
}

// The step of (mode, tau, post) captured into one executable graph (made on
// first use and kept until the plan is destroyed or another step is asked
// for).  The first run of a step is eager, so that every lazily made table
// and launch shape exists before the capture.
static int step_graph(DevicePlan *p, int mode, double tau, bool post, std::string &err) {
    DeviceCtx *c = p->ctx;
    if (p->graph && p->graph_mode == mode && p->graph_tau == tau && p->graph_post == post) return 0;
    int rc0 = drop_graph(p, err);
    if (rc0) return rc0;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(p, mode, tau, post, false, err);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        err = std::string("graph capture: ") + hipGetErrorString(e);
        return -2;
    }
    e = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        p->graph = nullptr;
        err = std::string("graph instantiation: ") + hipGetErrorString(e);
        return -2;
    }
    p->graph_mode = mode;
    p->graph_tau = tau;
    p->graph_post = post;
    return 0;
}

int dev_plan_run(DevicePlan *p, int mode, double tau, double *kernel_ms, unsigned long long *term_apps,
                 unsigned long long *lp_apps, std::string &err, double *split_ms, bool post, double *post_ms,
                 bool graph) {
    DeviceCtx *c = p->ctx;
    HIP_TRY(hipSetDevice(c->device));
    if (post && !p->post_pairs) {
        err = "the plan has no post transforms";
        return -1;
    }
    if (post_ms) *post_ms = 0.0;
    if (split_ms) split_ms[0] = split_ms[1] = 0.0;
    if (kernel_ms) *kernel_ms = 0.0;
    if (term_apps) *term_apps = 0;
    if (lp_apps) *lp_apps = 0;
    if (p->n_out == 0) return 0;
    if (mode < 0 || mode > 2) {
        err = "unknown arithmetic mode";
        return -1;
    }
    if (mode != 0 && !(tau >= 1.0)) {
        err = "mixed-precision ratio must be >= 1";
        return -1;
    }
    if (mode != 0) {
        int rc = mode == 1 ? ensure_f32(p, err) : ensure_h16(p, err);
        if (!rc) rc = make_lists(p, mode, tau, err);
        if (rc) return rc;
    }
    const unsigned bit = 1u << (2 * mode + (post ? 1 : 0));
    const bool replay = graph && (p->ran_eager & bit);
    if (replay) {
        int rc = step_graph(p, mode, tau, post, err);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        HIP_TRY(hipGraphLaunch(p->graph, c->stream));
    } else {
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        int rc = enqueue_step(p, mode, tau, post, true, err);
        if (rc) return rc;
        p->ran_eager |= bit;
    }
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    unsigned long long cnt[8];
    HIP_TRY(hipMemcpyAsync(cnt, p->args.counters, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (kernel_ms) *kernel_ms = ms;
    // per-launch times of an eager run (a replayed graph gives the whole only)
    if (split_ms && !replay) {
        if (mode == 0) {
            float mc = 0.f;
            HIP_TRY(hipEventElapsedTime(&mc, c->ev0, c->evp));
            split_ms[0] = mc;
        } else {
            // device time of each class launch (in the split mode they overlap)
            float m1 = 0.f, m2 = 0.f;
            HIP_TRY(hipEventElapsedTime(&m1, c->ev2s, c->ev2e));
            HIP_TRY(hipEventElapsedTime(&m2, c->evl, c->evm));
            split_ms[0] = m1;
            split_ms[1] = m2;
        }
    }
    if (post && post_ms && !replay) {
        float mp = 0.f;
        HIP_TRY(hipEventElapsedTime(&mp, c->evp, c->ev1));
        *post_ms = mp;
    }
    if (term_apps) *term_apps = cnt[0];
    if (lp_apps) *lp_apps = cnt[1];
    return 0;
}

int dev_plan_download(DevicePlan *p, double *out_ns, std::string &err) {
    DeviceCtx *c = p->ctx;
    HIP_TRY(hipSetDevice(c->device));
    if (p->n_out == 0) return 0;
    return dev_copy_d2h(c, out_ns, p->args.out_ns, sizeof(double) * (size_t)p->n_out * 8 * p->kp1 * p->kp1 * p->kp1,
                        err);
}

int dev_plan_clone_output(DevicePlan *p, double **out, std::string &err) {
    DeviceCtx *c = p->ctx;
    HIP_TRY(hipSetDevice(c->device));
    const size_t n = (size_t)std::max(p->n_out, 1) * 8 * p->kp1 * p->kp1 * p->kp1;
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(double) * n));
    hipError_t e = hipMemcpyAsync(d, p->args.out_ns, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipFree(d);
        err = std::string("output copy: ") + hipGetErrorString(e);
        return -2;
    }
    *out = d;
    return 0;
}

double *dev_plan_take_output(DevicePlan *p) {
    double *o = p->args.out_ns;
    for (void *&b : p->bufs)
        if (b == o) b = nullptr;
    p->args.out_ns = nullptr;
    return o;
}

int dev_plan_download_range(DevicePlan *p, long long first, long long count, double *dst, std::string &err) {
    DeviceCtx *c = p->ctx;
    if (first < 0 || count < 0 || first + count > p->n_out) {
        err = "output node range outside the plan";
        return -1;
    }
    if (count == 0) return 0;
    HIP_TRY(hipSetDevice(c->device));
    const size_t blk = (size_t)8 * p->kp1 * p->kp1 * p->kp1;
    return dev_copy_d2h(c, dst, p->args.out_ns + (size_t)first * blk, sizeof(double) * (size_t)count * blk, err);
}

}  // namespace mwg
