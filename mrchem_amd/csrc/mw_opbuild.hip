// Device construction of the non-standard operator tables (gfx950).
//
// The same tables as the host build (mw_operator.cpp), for a batch of
// operators at once: HelmholtzVector rebuilds one HelmholtzOperator per
// orbital in every SCF iteration (src/scf_solver/HelmholtzVector.cpp:144),
// and PoissonOperator is rebuilt in the SAD guess and the surface forces
// (src/driver.cpp:1131,1283).  The kernel fit stays on the host (a 1D fit of
// tens of terms); every (operator, term, depth) item becomes one wavefront:
//
//  pass 1 (k_op_scan): the translation scan of the host build, l = 0, 1, ...,
//    with the s<-s blocks T^{d+1}_delta from Gaussian moments (the same
//    piecewise Gauss-Legendre quadrature), the NS components of +l and -l
//    through the two-scale filter, their norms and the same stopping and
//    storage rules; it yields the stored half-width L of the item.
//  pass 2 (k_op_fill): the kept blocks and norms of l in [-L, L], written
//    straight into the flat operator table of the device.
//
// Sums run in a different order than on the host, so values agree to a few
// ulps; the tables stay resident on the device (the host keeps L and the
// norms, which the apply planning reads).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
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

// One (operator, term, depth) item.
struct OpItem {
    double A1, pref;       // T blocks at depth d + 1: exponent alpha h^2 and prefactor h coef1d
    double peak, tau;      // scan end position, storage floor parameter
    long long Lscan;       // scan bound
    long long nrm_off;     // pass 1: offset (doubles) of the item's norms, cap x 8
    long long boff, noff;  // pass 2: offsets of the item's blocks and norms in its table
    double *blocks, *norms;  // pass 2: the operator's flat tables
    int cap, d;
};
struct OpScanOut {
    int L, nscan, overflow, pad_;
    double floor_s, floor_w;
};

// Basis constants shared by all items (device pointers).
struct OpConst {
    const double *C;   // [i][j][p], 2N values per (i, j): cross-correlation coefficients
    const double *F;   // [2N][2N] two-scale filter
    const double *qx, *qw;
    int nq;
};

// Wave-wide sum (all lanes get it).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Moments int_0^1 P~_p(u) exp(-A (u - c)^2) du, p < N2, of the normalised
// shifted Legendre polynomials: the piecewise Gauss-Legendre rule of
// gauss_moments01 (mw_math.cpp), points spread over the wave.
template <int N2>
__device__ void moments(double A, double c, const OpConst &k, double *out, int lane) {
    double a = 0.0, b = 1.0;
    bool empty = false;
    if (A > 0.0) {
        const double vstar = fmin(1.0, fmax(0.0, c));
        const double dstar = fabs(vstar - c);
        if (A * dstar * dstar > 745.0) empty = true;
        const double R = sqrt(dstar * dstar + 46.0 / A);
        a = fmax(0.0, c - R);
        b = fmin(1.0, c + R);
        if (b <= a) empty = true;
    }
    double acc[N2];
#pragma unroll
    for (int p = 0; p < N2; p++) acc[p] = 0.0;
    if (!empty) {
        const double maxdist = fmax(fabs(a - c), fabs(b - c));
        double delta = b - a;
        if (A > 0.0) {
            delta = fmin(delta, 4.0 / sqrt(A));
            if (A * maxdist > 0.0) delta = fmin(delta, 8.0 / (A * maxdist));
        }
        int npieces = max(1, (int)ceil((b - a) / delta - 1e-12));
        if (npieces > 4096) npieces = 4096;
        const double width = (b - a) / npieces;
        const int total = npieces * k.nq;
        for (int e = lane; e < total; e += 64) {
            const int q = e / k.nq, i = e % k.nq;
            const double u0 = a + q * width;
            const double v = u0 + width * k.qx[i];
            const double g = exp(-A * (v - c) * (v - c)) * k.qw[i] * width;
            if (g == 0.0) continue;
            const double t = 2.0 * v - 1.0;
            double p0 = 1.0, p1 = t;
            acc[0] += g * 1.0;
            if (N2 > 1) acc[1] += g * (sqrt(3.0) * t);
#pragma unroll
            for (int p = 1; p + 1 < N2; p++) {
                const double p2 = ((2.0 * p + 1.0) * t * p1 - p * p0) / (p + 1.0);
                p0 = p1;
                p1 = p2;
                acc[p + 1] += g * (sqrt(2.0 * (p + 1) + 1.0) * p2);
            }
        }
    }
#pragma unroll
    for (int p = 0; p < N2; p++) {
        const double s = wave_sum(acc[p]);
        if (lane == 0) out[p] = s;
    }
}

// Rolling state of the scan over translations l = 0, 1, ...: moments
// M(2l), M(2l+1), M(2l+2) and blocks T(2l-1), T(2l), T(2l+1) in LDS.
template <int N>
struct OpState {
    static constexpr int N2 = 2 * N, NN = N * N, MM = N2 * N2;
    double *mom;  // [3][N2]
    double *T;    // [3][NN]: T(2l-1), T(2l), T(2l+1)
    double *Mf, *tmp, *Rp, *Rm;  // [MM] each
};

// T(delta)[i][j] = pref sum_p C[i][j][p] M1[p] + (-1)^p C[j][i][p] M2[p], with
// M1 = M(delta), M2 = M(delta + 1) (t_block, mw_operator.cpp).
template <int N>
__device__ void t_block(const OpConst &k, double pref, const double *M1, const double *M2, double *T, int lane) {
    constexpr int N2 = 2 * N;
    for (int e = lane; e < N * N; e += 64) {
        const int i = e / N, j = e % N;
        const double *cij = k.C + (i * N + j) * N2;
        const double *cji = k.C + (j * N + i) * N2;
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < N2; p++) s += cij[p] * M1[p] + ((p & 1) ? -cji[p] : cji[p]) * M2[p];
        T[e] = pref * s;
    }
}

// R = F Mf F^T with Mf = [[A, B], [Cb, A]] built from three N x N blocks, each
// possibly transposed.
template <int N>
__device__ void ns_transform(const OpConst &k, const double *A, bool tA, const double *B, bool tB, const double *Cb,
                             bool tC, double *Mf, double *tmp, double *R, int lane) {
    constexpr int N2 = 2 * N, MM = N2 * N2;
    for (int e = lane; e < MM; e += 64) {
        const int r = e / N2, c = e % N2;
        const int bi = r % N, bj = c % N;
        const double *src = (r < N) ? (c < N ? A : B) : (c < N ? Cb : A);
        const bool tr = (r < N) ? (c < N ? tA : tB) : (c < N ? tC : tA);
        Mf[e] = tr ? src[bj * N + bi] : src[bi * N + bj];
    }
    __syncthreads();
    for (int e = lane; e < MM; e += 64) {
        const int i = e / N2, j = e % N2;
        double s = 0.0;
#pragma unroll 4
        for (int q = 0; q < N2; q++) s += k.F[i * N2 + q] * Mf[q * N2 + j];
        tmp[e] = s;
    }
    __syncthreads();
    for (int e = lane; e < MM; e += 64) {
        const int i = e / N2, j = e % N2;
        double s = 0.0;
#pragma unroll 4
        for (int q = 0; q < N2; q++) s += tmp[i * N2 + q] * k.F[j * N2 + q];
        R[e] = s;
    }
    __syncthreads();
}

// Frobenius norms of the four components c = 2 gt + ft of R (row half gt,
// column half ft).
template <int N>
__device__ void comp_norms(const double *R, double *nr, int lane) {
    constexpr int N2 = 2 * N;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int r0 = (c >> 1) * N, c0 = (c & 1) * N;
        double s = 0.0;
        for (int e = lane; e < N * N; e += 64) {
            const double v = R[(r0 + e / N) * N2 + c0 + e % N];
            s += v * v;
        }
        nr[c] = sqrt(wave_sum(s));
    }
}

// Advance the rolling state to translation l and form R+ = NS(l), R- = NS(-l).
template <int N>
__device__ void ns_step(const OpConst &k, const OpItem &it, OpState<N> &st, long long l, int lane) {
    constexpr int N2 = 2 * N, NN = N * N;
    double *M0 = st.mom, *M1 = st.mom + N2, *M2 = st.mom + 2 * N2;
    double *Tm = st.T, *T0 = st.T + NN, *Tp = st.T + 2 * NN;
    if (l == 0) {
        moments<N2>(it.A1, 0.0, k, M0, lane);
    } else {
        // M(2l) = previous M(2l), stored in slot 2; T(2l - 1) = previous T(2l + 1)
        __syncthreads();
        for (int e = lane; e < N2; e += 64) M0[e] = M2[e];
        for (int e = lane; e < NN; e += 64) Tm[e] = Tp[e];
    }
    moments<N2>(it.A1, (double)(2 * l + 1), k, M1, lane);
    moments<N2>(it.A1, (double)(2 * l + 2), k, M2, lane);
    __syncthreads();
    t_block<N>(k, it.pref, M0, M1, T0, lane);
    t_block<N>(k, it.pref, M1, M2, Tp, lane);
    __syncthreads();
    if (l == 0) {
        // T(-1) = T(1)^T
        for (int e = lane; e < NN; e += 64) Tm[e] = Tp[(e % N) * N + e / N];
        __syncthreads();
    }
    // NS(l): [[T(2l), T(2l+1)], [T(2l-1), T(2l)]]
    ns_transform<N>(k, T0, false, Tp, false, Tm, false, st.Mf, st.tmp, st.Rp, lane);
    // NS(-l): [[T(-2l), T(-2l+1)], [T(-2l-1), T(-2l)]] = transposes of T(2l), T(2l-1), T(2l+1)
    ns_transform<N>(k, T0, true, Tm, true, Tp, true, st.Mf, st.tmp, st.Rm, lane);
}

template <int N>
__device__ void setup_state(OpState<N> &st, double *lds) {
    constexpr int N2 = 2 * N, NN = N * N, MM = N2 * N2;
    st.mom = lds;
    st.T = st.mom + 3 * N2;
    st.Mf = st.T + 3 * NN;
    st.tmp = st.Mf + MM;
    st.Rp = st.tmp + MM;
    st.Rm = st.Rp + MM;
}
template <int N>
constexpr int op_lds_doubles() {
    return 3 * 2 * N + 3 * N * N + 4 * (2 * N) * (2 * N);
}

// Pass 1: the translation scan of build_term_depth (mw_operator.cpp): the
// norms of every scanned translation, the scan length and the storage floors.
template <int N>
__global__ __launch_bounds__(64) void k_op_scan(const OpItem *items, OpConst k, double *nrm, OpScanOut *out) {
    __shared__ double lds[op_lds_doubles<N>()];
    const int lane = threadIdx.x;
    const OpItem it = items[blockIdx.x];
    OpState<N> st;
    setup_state<N>(st, lds);
    double ref = 0.0, wmax = 0.0;
    int quiet = 0;
    long long l = 0, nscan = 0;
    int overflow = 0;
    for (; l <= it.Lscan; l++) {
        if (l >= it.cap) {
            overflow = 1;
            break;
        }
        ns_step<N>(k, it, st, l, lane);
        double a[4], b[4];
        comp_norms<N>(st.Rp, a, lane);
        comp_norms<N>(st.Rm, b, lane);
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                nrm[it.nrm_off + l * 8 + c] = a[c];
                nrm[it.nrm_off + l * 8 + 4 + c] = b[c];
            }
        }
        nscan = l + 1;
        double wrow = 0.0, srow = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            ref = fmax(ref, fmax(a[c], b[c]));
            if (c == 0) srow = fmax(a[c], b[c]);
            else wrow = fmax(wrow, fmax(a[c], b[c]));
        }
        wmax = fmax(wmax, fmax(wrow, 0.0));
        const double fs = (it.d == 0) ? it.tau * ref : it.tau * ref * ref / fmax(wmax, 1e-300);
        if (wrow > it.tau * ref || srow > fs) quiet = 0;
        else quiet++;
        if (quiet >= 3 && (double)l >= it.peak) break;
    }
    const double floor_w = it.tau * ref;
    const double floor_s = (it.d == 0) ? it.tau * ref : it.tau * ref * ref / fmax(wmax, 1e-300);
    if (lane == 0) {
        OpScanOut r;
        r.L = -1;  // set by the host from the norms
        r.nscan = (int)nscan;
        r.overflow = overflow;
        r.pad_ = 0;
        r.floor_s = floor_s;
        r.floor_w = floor_w;
        out[blockIdx.x] = r;
    }
}

// Pass 2: blocks and norms of l in [-L, L] into the flat table [c][l + L][i][j];
// components below their floor stay zero (the table is cleared beforehand).
template <int N>
__global__ __launch_bounds__(64) void k_op_fill(const OpItem *items, OpConst k, const OpScanOut *scan) {
    __shared__ double lds[op_lds_doubles<N>()];
    constexpr int N2 = 2 * N, NN = N * N;
    const int lane = threadIdx.x;
    const OpItem it = items[blockIdx.x];
    const OpScanOut so = scan[blockIdx.x];
    if (so.L < 0) return;
    const int L = so.L, nl = 2 * L + 1;
    OpState<N> st;
    setup_state<N>(st, lds);
    for (long long l = 0; l <= L; l++) {
        ns_step<N>(k, it, st, l, lane);
        double a[4], b[4];
        comp_norms<N>(st.Rp, a, lane);
        comp_norms<N>(st.Rm, b, lane);
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const double *R = side == 0 ? st.Rp : st.Rm;
            const long long ls = side == 0 ? l : -l;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const double nr = side == 0 ? a[c] : b[c];
                const bool keep = nr > (c == 0 ? so.floor_s : so.floor_w);
                if (!keep) continue;
                const long long idx = (long long)c * nl + (ls + L);
                const int r0 = (c >> 1) * N, c0 = (c & 1) * N;
                for (int e = lane; e < NN; e += 64)
                    it.blocks[it.boff + idx * NN + e] = R[(r0 + e / N) * N2 + c0 + e % N];
                if (lane == 0) it.norms[it.noff + idx] = nr;
            }
        }
    }
}

namespace {

template <int N>
int launch_scan(const OpItem *items, int n, const OpConst &k, double *nrm, OpScanOut *out, hipStream_t s,
                std::string &err) {
    if (n == 0) return 0;
    hipLaunchKernelGGL((k_op_scan<N>), dim3(n), dim3(64), 0, s, items, k, nrm, out);
    HIP_TRY(hipGetLastError());
    return 0;
}
template <int N>
int launch_fill(const OpItem *items, int n, const OpConst &k, const OpScanOut *scan, hipStream_t s, std::string &err) {
    if (n == 0) return 0;
    hipLaunchKernelGGL((k_op_fill<N>), dim3(n), dim3(64), 0, s, items, k, scan);
    HIP_TRY(hipGetLastError());
    return 0;
}

#define MWG_OP_SWITCH(N, CALL)                                  \
    switch (N) {                                                \
        case 2: return CALL(2);                                 \
        case 3: return CALL(3);                                 \
        case 4: return CALL(4);                                 \
        case 5: return CALL(5);                                 \
        case 6: return CALL(6);                                 \
        case 7: return CALL(7);                                 \
        case 8: return CALL(8);                                 \
        case 9: return CALL(9);                                 \
        case 10: return CALL(10);                               \
        case 11: return CALL(11);                               \
        case 12: return CALL(12);                               \
        default: err = "unsupported basis size"; return -1;     \
    }

int scan_n(int N, const OpItem *items, int n, const OpConst &k, double *nrm, OpScanOut *out, hipStream_t s,
           std::string &err) {
#define MWG_CALL(NN) launch_scan<NN>(items, n, k, nrm, out, s, err)
    MWG_OP_SWITCH(N, MWG_CALL)
#undef MWG_CALL
}
int fill_n(int N, const OpItem *items, int n, const OpConst &k, const OpScanOut *scan, hipStream_t s,
           std::string &err) {
#define MWG_CALL(NN) launch_fill<NN>(items, n, k, scan, s, err)
    MWG_OP_SWITCH(N, MWG_CALL)
#undef MWG_CALL
}

// Device buffers freed on scope exit.
struct DevBufs {
    std::vector<void *> p;
    template <typename T>
    int alloc(size_t n, T **out, std::string &err) {
        void *d = nullptr;
        HIP_TRY(hipMalloc(&d, sizeof(T) * std::max<size_t>(n, 1)));
        p.push_back(d);
        *out = static_cast<T *>(d);
        return 0;
    }
    ~DevBufs() {
        for (void *q : p) (void)hipFree(q);
    }
};

}  // namespace

// Translations a scan may hold before the build reports an error.
constexpr long long kOpScanCap = 4096;

int dev_build_operators(DeviceCtx *c, const MRA &mra, const std::vector<OpBuildJob> &jobs, std::string &err) {
    HIP_TRY(hipSetDevice(dev_device(c)));
    hipStream_t s = static_cast<hipStream_t>(dev_stream(c));
    const int N = mra.kp1(), N2 = 2 * N, NN = N * N;
    if (N < 2 || N > 12) {
        err = "unsupported basis size k+1 = " + std::to_string(N);
        return -1;
    }
    // items of every (operator, term, depth) that stores blocks
    std::vector<OpItem> items;
    std::vector<int> item_job, item_m;
    long long nrm_total = 0;
    for (size_t jb = 0; jb < jobs.size(); jb++) {
        const Operator &op = *jobs[jb].op;
        const double tau = operator_tau(op.build_prec);
        for (int m = 0; m < op.size(); m++) {
            for (int d = 0; d <= op.max_depth; d++) {
                long long Lscan = 0;
                double peak = 0.0;
                if (!term_depth_scan(mra, op.terms[m].alpha, d, tau, Lscan, peak)) continue;
                OpItem it{};
                const double h1 = mra.box_size(d + 1);
                it.A1 = op.terms[m].alpha * h1 * h1;
                it.pref = h1 * op.terms[m].coef1d;
                it.peak = peak;
                it.tau = tau;
                it.Lscan = Lscan;
                it.cap = (int)std::min(Lscan + 1, kOpScanCap);
                it.d = d;
                it.nrm_off = nrm_total;
                nrm_total += 8LL * it.cap;
                items.push_back(it);
                item_job.push_back((int)jb);
                item_m.push_back(m);
            }
        }
    }
    if (items.size() >= (size_t)INT_MAX) {
        err = "too many operator terms for one device build";
        return -1;
    }
    const int n_items = (int)items.size();
    // basis constants
    const double *qx = nullptr, *qw = nullptr;
    const int nq = N2 + 20;
    quad_rule01(nq, &qx, &qw);
    DevBufs tmp;
    double *dC, *dF, *dqx, *dqw, *dnrm;
    OpItem *ditems;
    OpScanOut *dscan;
    int rc = tmp.alloc(mra.basis.C.size(), &dC, err);
    if (!rc) rc = tmp.alloc(mra.basis.F.size(), &dF, err);
    if (!rc) rc = tmp.alloc(nq, &dqx, err);
    if (!rc) rc = tmp.alloc(nq, &dqw, err);
    if (!rc) rc = tmp.alloc((size_t)nrm_total, &dnrm, err);
    if (!rc) rc = tmp.alloc(items.size(), &ditems, err);
    if (!rc) rc = tmp.alloc(items.size(), &dscan, err);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(dC, mra.basis.C.data(), sizeof(double) * mra.basis.C.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dF, mra.basis.F.data(), sizeof(double) * mra.basis.F.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dqx, qx, sizeof(double) * nq, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dqw, qw, sizeof(double) * nq, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ditems, items.data(), sizeof(OpItem) * items.size(), hipMemcpyHostToDevice, s));
    OpConst k{dC, dF, dqx, dqw, nq};
    rc = scan_n(N, ditems, n_items, k, dnrm, dscan, s, err);
    if (rc) return rc;
    std::vector<OpScanOut> scan(items.size());
    HIP_TRY(hipMemcpyAsync(scan.data(), dscan, sizeof(OpScanOut) * scan.size(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<double> nrm_h((size_t)nrm_total);
    HIP_TRY(hipMemcpy(nrm_h.data(), dnrm, sizeof(double) * nrm_h.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < n_items; i++) {
        OpScanOut &so = scan[i];
        if (so.overflow) {
            err = "operator translation scan exceeded " + std::to_string(kOpScanCap) + " translations";
            return -1;
        }
        // stored half-width: the largest scanned translation with a component
        // above its floor (build_term_depth)
        const double *nr = nrm_h.data() + items[i].nrm_off;
        so.L = -1;
        for (int l = 0; l < so.nscan; l++)
            for (int c = 0; c < 4; c++) {
                const double fl = c == 0 ? so.floor_s : so.floor_w;
                if (nr[l * 8 + c] > fl || nr[l * 8 + 4 + c] > fl) so.L = l;
            }
    }
    HIP_TRY(hipMemcpyAsync(dscan, scan.data(), sizeof(OpScanOut) * scan.size(), hipMemcpyHostToDevice, s));
    // flat layout of each operator (the order of flatten_operator)
    std::vector<std::vector<int>> Lmd(jobs.size());
    for (size_t jb = 0; jb < jobs.size(); jb++) {
        const Operator &op = *jobs[jb].op;
        Lmd[jb].assign((size_t)op.size() * (op.max_depth + 1), -1);
    }
    for (int i = 0; i < n_items; i++) {
        const Operator &op = *jobs[item_job[i]].op;
        Lmd[item_job[i]][(size_t)item_m[i] * (op.max_depth + 1) + items[i].d] = scan[i].L;
    }
    std::vector<DevOperTables> tabs(jobs.size());
    for (size_t jb = 0; jb < jobs.size(); jb++) {
        const Operator &op = *jobs[jb].op;
        OperFlat &f = *jobs[jb].flat;
        f = OperFlat();
        f.M = op.size();
        f.ndepth = op.max_depth + 1;
        f.kp1 = N;
        f.L = Lmd[jb];
        f.boff.assign(f.L.size(), 0);
        f.noff.assign(f.L.size(), 0);
        long long nb = 0;
        for (size_t i = 0; i < f.L.size(); i++) {
            f.boff[i] = nb * NN;
            f.noff[i] = nb;
            if (f.L[i] >= 0) nb += 4LL * (2 * f.L[i] + 1);
        }
        f.norms.assign((size_t)nb, 0.0);
        f.device_blocks = nb * NN;
        rc = dev_alloc_oper_tables(c, f, tabs[jb], err);
        if (rc) return rc;
    }
    for (int i = 0; i < n_items; i++) {
        const OperFlat &f = *jobs[item_job[i]].flat;
        const size_t md = (size_t)item_m[i] * f.ndepth + items[i].d;
        items[i].boff = f.boff[md];
        items[i].noff = f.noff[md];
        items[i].blocks = tabs[item_job[i]].blocks;
        items[i].norms = tabs[item_job[i]].norms;
    }
    HIP_TRY(hipMemcpyAsync(ditems, items.data(), sizeof(OpItem) * items.size(), hipMemcpyHostToDevice, s));
    rc = fill_n(N, ditems, n_items, k, dscan, s, err);
    if (rc) return rc;
    // norms back to the host (the planning reads them)
    for (size_t jb = 0; jb < jobs.size(); jb++) {
        OperFlat &f = *jobs[jb].flat;
        if (!f.norms.empty())
            HIP_TRY(hipMemcpyAsync(f.norms.data(), tabs[jb].norms, sizeof(double) * f.norms.size(),
                                   hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    for (size_t jb = 0; jb < jobs.size(); jb++) {
        Operator &op = *jobs[jb].op;
        const OperFlat &f = *jobs[jb].flat;
        for (int m = 0; m < op.size(); m++)
            for (int d = 0; d <= op.max_depth; d++) {
                const size_t md = (size_t)m * f.ndepth + d;
                OperDepth &od = op.terms[m].depth[d];
                od.L = f.L[md];
                od.blocks.clear();
                od.norms.clear();
                if (od.L >= 0)
                    od.norms.assign(f.norms.begin() + f.noff[md], f.norms.begin() + f.noff[md] + 4 * (2 * od.L + 1));
            }
        rc = dev_register_operator(c, jobs[jb].key, f, tabs[jb], err);
        if (rc) return rc;
    }
    return 0;
}

}  // namespace mwg
