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
#include <map>
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

constexpr int kMaxActive = 4096;

struct KArgs {
    // output nodes
    const int *out_depth;
    const double *out_thr;
    const long long *band_off;
    const int *band_w;
    const int *band;
    double *out_ns;
    // input slots
    const long long *slot_off;
    const double *pool;
    const double *slot_norm;
    // operator
    const int *opL;
    const long long *op_boff, *op_noff;
    const double *op_blocks, *op_norms;
    int M, ndepth;
    // screening tables
    const int *bw;
    const int *act_off, *act_terms;
    // [0] term-applications passing screening (algorithmic work)
    unsigned long long *counters;
};

__device__ __forceinline__ double op_norm(const KArgs &a, int m, int d, int c, int l) {
    int L = a.opL[m * a.ndepth + d];
    if (l < -L || l > L) return 0.0;
    return a.op_norms[a.op_noff[m * a.ndepth + d] + (long long)c * (2 * L + 1) + (l + L)];
}

// One workgroup per job (output node g, output component gt); its NW waves
// share the job's 16-wide tiles (TW tiles each).  The active terms of the
// job's depth are taken in chunks of 64; for the chunk, the 1D norms of every
// term, component and translation in the band are staged in LDS, so that the
// screening of a band node f reads six LDS values per lane: the bound
// ||O_x|| ||O_y|| ||O_z|| ||f_ft|| for every input component ft, one ballot per
// ft.  Every wave repeats the screening and holds the same ballot.  Each
// surviving (ft, term) is applied as three chains of v_mfma_f64_16x16x4:
//   x: X[(jz,jy)][ix]  = F[(jz,jy)][jx] . Ox^T[jx][ix]   F tiles in registers
//   y: Y[iy][(jz,ix)]  = Oy[iy][jy] . X'[jy][(jz,ix)]
//   z: acc[iz][(iy,ix)] += Oz[iz][jz] . Y''[jz][(iy,ix)]  acc tiles persist
// X and Y change layout through one padded LDS buffer whose every lane
// address is a base plus a compile-time offset (padding entries stay zero or
// finite and only meet zero operands).  The three operator blocks of the next
// term are loaded into a double-buffered LDS area while the current term
// computes.  MFMA lane layout (checked on gfx950): A[i][k] in lane i + 16k,
// B[k][j] in lane j + 16k, D[i][j] in lane j + 16 (i mod 4), register i / 4.
typedef double mwg_double4 __attribute__((ext_vector_type(4)));

template <int N>
struct JobShape {
    static constexpr int N2 = N * N, N3 = N2 * N;
    static constexpr int NT = (N2 + 15) / 16;  // 16-wide tiles over N^2
    static constexpr int KS = (N + 3) / 4;     // 4-deep k steps over N
    static constexpr int TW = 4;               // tiles per wave
    static constexpr int NW = (NT + TW - 1) / TW;
    static constexpr int T = 64 * NW;          // threads per job
    static constexpr int NTP = NW * TW;        // tiles including padding tiles
    // Xp[row][16], row = (jz,jy) flattened; Yp[16][CP], [iy][(jz,ix)]
    static constexpr int XROWS_W = 16 * NTP;
    static constexpr int XROWS_R = ((16 * NTP - 1) / N) * N + 4 * KS;
    static constexpr int XROWS = XROWS_W > XROWS_R ? XROWS_W : XROWS_R;
    static constexpr int CP0 = 16 * NTP > 4 * KS * N ? 16 * NTP : 4 * KS * N;
    static constexpr int CP = CP0 + ((16 * NTP - 1) / N >= 16 ? 0 : 0);
    static constexpr int XSZ = XROWS * 16, YSZ = 16 * CP;
    static constexpr int BUF = XSZ > YSZ ? XSZ : YSZ;
    static constexpr int OB = 3 * N2;          // operator doubles per stage
    static constexpr int OPL = (OB + T - 1) / T;
};

constexpr int kMaxBand = 9;  // band half-width limit of the LDS norm stage

template <int N>
__global__ __launch_bounds__(JobShape<N>::T) void k_apply_fp64(KArgs a) {
    using S = JobShape<N>;
    constexpr int N2 = S::N2, N3 = S::N3, KS = S::KS, TW = S::TW, T = S::T;
    constexpr int CP = S::CP, BUF = S::BUF, OB = S::OB, OPL = S::OPL;
    __shared__ double buf[BUF];
    __shared__ double ops[2][OB];
    extern __shared__ double nrm_s[];  // [term][c][l+W], 64 x 4 x (2 Wmax + 1)
    __shared__ int ltab[64];
    __shared__ long long btab[64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lc = lane & 15, lq = lane >> 4;
    for (int e = tid; e < BUF; e += T) buf[e] = 0.0;

    const int job = blockIdx.x;
    const int g = job >> 3, gt = job & 7;
    const int d = a.out_depth[g];
    const double thr = a.out_thr[g];
    const int w = a.band_w[g], side = 2 * w + 1;
    const long long boff = a.band_off[g];
    const int a0 = a.act_off[d], nact = a.act_off[d + 1] - a0;
    const int ndepth = a.ndepth;
    const int gbx = gt & 1, gby = (gt >> 1) & 1, gbz = (gt >> 2) & 1;

    // this wave's tiles: t = wv*TW + tt
    int ybase[TW], zbase[TW];
#pragma unroll
    for (int tt = 0; tt < TW; tt++) {
        const int col = 16 * (wv * TW + tt) + lc;
        const int q = col / N, r = col % N;
        ybase[tt] = (q * N + lq) * 16 + r;  // Xp[(jz*N + 4k + lq)][ix]
        zbase[tt] = q * CP + r + lq * N;    // Yp[iy][(4k + lq)*N + ix]
    }
    mwg_double4 acc[TW];
#pragma unroll
    for (int tt = 0; tt < TW; tt++) acc[tt] = mwg_double4{0.0, 0.0, 0.0, 0.0};
    double fa[TW][KS];  // x-pass A operand: F[16t + lc][4k + lq]
    unsigned long long napps = 0;
    const int nband = side * side * side;
    const int nl = 2 * w + 1;

    for (int base = 0; base < nact; base += 64) {
        const int cnt = min(64, nact - base);
        __syncthreads();
        // stage the chunk's table metadata and band norms
        for (int e = tid; e < 64; e += T) {
            const int m = (e < cnt) ? a.act_terms[a0 + base + e] : 0;
            ltab[e] = (e < cnt) ? a.opL[m * ndepth + d] : -1;
            btab[e] = (e < cnt) ? a.op_boff[m * ndepth + d] : 0;
        }
        for (int e = tid; e < 64 * 4 * nl; e += T) {
            const int ti = e / (4 * nl), c = (e / nl) % 4, li = e % nl, l = li - w;
            double v = 0.0;
            if (ti < cnt) {
                const int m = a.act_terms[a0 + base + ti];
                const int L = a.opL[m * ndepth + d];
                const int bwc = a.bw[((long long)d * a.M + m) * 4 + c];
                if (abs(l) <= bwc && abs(l) <= L)
                    v = a.op_norms[a.op_noff[m * ndepth + d] + (long long)c * (2 * L + 1) + (l + L)];
            }
            nrm_s[e] = v;
        }
        __syncthreads();
        const bool have = lane < cnt;
        for (int p = 0; p < nband; p++) {
            const int s = a.band[boff + p];
            if (s < 0) continue;
            const int ll[3] = {p % side - w, (p / side) % side - w, p / (side * side) - w};
            double nrm[2][3];
            const int gb[3] = {gbx, gby, gbz};
#pragma unroll
            for (int dim = 0; dim < 3; dim++)
#pragma unroll
                for (int fb = 0; fb < 2; fb++)
                    nrm[fb][dim] = nrm_s[(lane * 4 + (2 * gb[dim] + fb)) * nl + (ll[dim] + w)];
            const double *fnp = a.slot_norm + (long long)s * 8;
            for (int ft = 0; ft < 8; ft++) {
                if (d > 0 && ft == 0 && gt == 0) continue;
                const double fn = fnp[ft];
                if (!(fn > 0.0)) continue;
                const int fx = ft & 1, fy = (ft >> 1) & 1, fz = (ft >> 2) & 1;
                const double bound = ((nrm[fx][0] * nrm[fy][1]) * nrm[fz][2]) * fn;
                unsigned long long mask = __ballot(have && bound > thr);
                if (!mask) continue;
                {
                    const double *fsrc = a.pool + a.slot_off[s] + (long long)ft * N3;
#pragma unroll
                    for (int tt = 0; tt < TW; tt++) {
                        const int row = 16 * (wv * TW + tt) + lc;
#pragma unroll
                        for (int k = 0; k < KS; k++) {
                            const int jx = 4 * k + lq;
                            fa[tt][k] = (row < N2 && jx < N) ? fsrc[row * N + jx] : 0.0;
                        }
                    }
                }
                const int cxyz[3] = {(gbx << 1) | fx, (gby << 1) | fy, (gbz << 1) | fz};
                auto fetch = [&](int b, double *regs) {
                    const int Lb = ltab[b];
                    const long long bb = btab[b];
#pragma unroll
                    for (int u = 0; u < OPL; u++) {
                        const int e = tid + T * u;
                        regs[u] = 0.0;
                        if (e < OB) {
                            const int dim = e / N2, r = e % N2;
                            regs[u] = a.op_blocks[bb + ((long long)cxyz[dim] * (2 * Lb + 1) + (ll[dim] + Lb)) * N2 + r];
                        }
                    }
                };
                double pre[OPL];
                int b = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                fetch(b, pre);
                int stage = 0;
                while (true) {
#pragma unroll
                    for (int u = 0; u < OPL; u++) {
                        const int e = tid + T * u;
                        if (e < OB) ops[stage][e] = pre[u];
                    }
                    const bool more = mask != 0;
                    if (more) {
                        b = __ffsll((long long)mask) - 1;
                        mask &= mask - 1;
                        fetch(b, pre);  // next term's blocks in flight during this one
                    }
                    __syncthreads();
                    napps++;
                    // MFMA operands of the operators: Ox[lc][4k+lq] (B of x,
                    // as Ox^T), Oy[lc][4k+lq], Oz[lc][4k+lq] (A of y, z)
                    double ox[KS], oy[KS], oz[KS];
                    const double *O = ops[stage];
#pragma unroll
                    for (int k = 0; k < KS; k++) {
                        const int j = 4 * k + lq;
                        const bool v = lc < N && j < N;
                        ox[k] = v ? O[lc * N + j] : 0.0;
                        oy[k] = v ? O[N2 + lc * N + j] : 0.0;
                        oz[k] = v ? O[2 * N2 + lc * N + j] : 0.0;
                    }
                    // x pass -> Xp[row][ix], row = 16t + lq + 4r, ix = lc
                    {
                        mwg_double4 dx[TW];
#pragma unroll
                        for (int tt = 0; tt < TW; tt++) dx[tt] = mwg_double4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                        for (int k = 0; k < KS; k++)
#pragma unroll
                            for (int tt = 0; tt < TW; tt++)
                                dx[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[tt][k], ox[k], dx[tt], 0, 0, 0);
#pragma unroll
                        for (int tt = 0; tt < TW; tt++)
#pragma unroll
                            for (int r = 0; r < 4; r++)
                                buf[(lq * 16 + lc) + (16 * (wv * TW + tt) + 4 * r) * 16] = dx[tt][r];
                    }
                    __syncthreads();
                    // y pass: B[jy][col] = X[jz*N + jy][ix] -> Yp[iy][col]
                    {
                        mwg_double4 dy[TW];
#pragma unroll
                        for (int tt = 0; tt < TW; tt++) dy[tt] = mwg_double4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                        for (int k = 0; k < KS; k++)
#pragma unroll
                            for (int tt = 0; tt < TW; tt++)
                                dy[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(oy[k], buf[ybase[tt] + 64 * k], dy[tt], 0, 0, 0);
                        __syncthreads();
#pragma unroll
                        for (int tt = 0; tt < TW; tt++)
#pragma unroll
                            for (int r = 0; r < 4; r++)
                                buf[(lq * CP + lc) + (4 * r * CP + 16 * (wv * TW + tt))] = dy[tt][r];
                    }
                    __syncthreads();
                    // z pass: B[jz][col'] = Yp[iy][jz*N + ix], col' = iy*N + ix
#pragma unroll
                    for (int k = 0; k < KS; k++)
#pragma unroll
                        for (int tt = 0; tt < TW; tt++)
                            acc[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(oz[k], buf[zbase[tt] + 4 * k * N], acc[tt], 0, 0, 0);
                    if (!more) break;
                    stage ^= 1;
                }
                __syncthreads();
            }
        }
    }
    if (tid == 0 && napps) atomicAdd(&a.counters[0], napps);
    // acc[tt][r] = out[iz = lq + 4r][col = 16t + lc]
    double *dst = a.out_ns + (long long)g * 8 * N3 + (long long)gt * N3;
#pragma unroll
    for (int tt = 0; tt < TW; tt++) {
        const int col = 16 * (wv * TW + tt) + lc;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int iz = lq + 4 * r;
            if (col < N2 && iz < N) dst[iz * N2 + col] = acc[tt][r];
        }
    }
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
};

struct DeviceCtx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::map<const void *, DevOper> opers;
};

struct DevicePlan {
    DeviceCtx *ctx = nullptr;
    int n_out = 0, kp1 = 0, wmax = 0;
    KArgs args{};
    std::vector<void *> bufs;
    size_t out_bytes = 0;
    ~DevicePlan() {
        for (void *p : bufs)
            if (p) (void)hipFree(p);
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
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        err = std::string("stream/event creation: ") + hipGetErrorString(e);
        delete c;
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
    o = DevOper();
}

void dev_destroy(DeviceCtx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->opers) free_oper(kv.second);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void dev_forget_operator(DeviceCtx *c, const void *key) {
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

static int get_oper(DeviceCtx *c, const void *key, const OperFlat &f, DevOper **out, std::string &err) {
    auto it = c->opers.find(key);
    if (it != c->opers.end()) {
        *out = &it->second;
        return 0;
    }
    DevOper o;
    int rc = alloc_copy(f.L, &o.L, c->stream, err);
    if (!rc) rc = alloc_copy(f.boff, &o.boff, c->stream, err);
    if (!rc) rc = alloc_copy(f.noff, &o.noff, c->stream, err);
    if (!rc) rc = alloc_copy(f.blocks, &o.blocks, c->stream, err);
    if (!rc) rc = alloc_copy(f.norms, &o.norms, c->stream, err);
    if (rc) {
        free_oper(o);
        return rc;
    }
    c->opers[key] = o;
    *out = &c->opers[key];
    return 0;
}

template <typename T>
static int plan_copy(DevicePlan *p, const std::vector<T> &v, const T **out, std::string &err) {
    T *d = nullptr;
    int rc = alloc_copy(v, &d, p->ctx->stream, err);
    p->bufs.push_back(d);
    if (rc) return rc;
    *out = d;
    return 0;
}

int dev_plan_create(DeviceCtx *c, const void *op_key, const OperFlat &f, const ApplyPlan &plan, DevicePlan **out,
                    std::string &err) {
    HIP_TRY(hipSetDevice(c->device));
    for (size_t d = 0; d + 1 < plan.active_off.size(); d++)
        if (plan.active_off[d + 1] - plan.active_off[d] > kMaxActive) {
            err = "more than " + std::to_string(kMaxActive) + " active terms at one depth";
            return -1;
        }
    for (int wq : plan.band_w)
        if (wq > kMaxBand) {
            err = "band half-width " + std::to_string(wq) + " exceeds the kernel limit " + std::to_string(kMaxBand);
            return -1;
        }
    if (plan.kp1 < 2 || plan.kp1 > 12) {
        err = "unsupported basis size k+1 = " + std::to_string(plan.kp1);
        return -1;
    }
    DevOper *o = nullptr;
    int rc = get_oper(c, op_key, f, &o, err);
    if (rc) return rc;
    auto *p = new DevicePlan;
    p->ctx = c;
    p->n_out = (int)plan.out_node.size();
    p->kp1 = plan.kp1;
    for (int wq : plan.band_w) p->wmax = std::max(p->wmax, wq);
    KArgs &a = p->args;
    rc = plan_copy(p, plan.out_depth, &a.out_depth, err);
    if (!rc) rc = plan_copy(p, plan.out_thr, &a.out_thr, err);
    if (!rc) rc = plan_copy(p, plan.band_off, &a.band_off, err);
    if (!rc) rc = plan_copy(p, plan.band_w, &a.band_w, err);
    if (!rc) rc = plan_copy(p, plan.band, &a.band, err);
    if (!rc) rc = plan_copy(p, plan.slot_off, &a.slot_off, err);
    if (!rc) rc = plan_copy(p, plan.in_coefs, &a.pool, err);
    if (!rc) rc = plan_copy(p, plan.in_norms, &a.slot_norm, err);
    if (!rc) rc = plan_copy(p, plan.bw, &a.bw, err);
    if (!rc) rc = plan_copy(p, plan.active_off, &a.act_off, err);
    if (!rc) rc = plan_copy(p, plan.active_terms, &a.act_terms, err);
    if (rc) {
        delete p;
        return rc;
    }
    const int N = plan.kp1;
    p->out_bytes = sizeof(double) * (size_t)std::max(p->n_out, 1) * 8 * N * N * N;
    void *d_out = nullptr, *d_cnt = nullptr;
    hipError_t e = hipMalloc(&d_out, p->out_bytes);
    if (e == hipSuccess) p->bufs.push_back(d_out);
    if (e == hipSuccess) e = hipMalloc(&d_cnt, 64);
    if (e == hipSuccess) p->bufs.push_back(d_cnt);
    if (e != hipSuccess) {
        err = std::string("hipMalloc: ") + hipGetErrorString(e);
        delete p;
        return -2;
    }
    a.out_ns = static_cast<double *>(d_out);
    a.counters = static_cast<unsigned long long *>(d_cnt);
    a.opL = o->L;
    a.op_boff = o->boff;
    a.op_noff = o->noff;
    a.op_blocks = o->blocks;
    a.op_norms = o->norms;
    a.M = f.M;
    a.ndepth = f.ndepth;
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        err = std::string("upload: ") + hipGetErrorString(e);
        delete p;
        return -2;
    }
    *out = p;
    return 0;
}

void dev_plan_destroy(DevicePlan *p) {
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    delete p;
}

template <int N>
static int launch_n(const KArgs &a, int n_out, int wmax, hipStream_t s, std::string &err) {
    const size_t shmem = sizeof(double) * 64 * 4 * (2 * wmax + 1);
    hipLaunchKernelGGL(k_apply_fp64<N>, dim3(8 * n_out), dim3(JobShape<N>::T), shmem, s, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

int dev_plan_run(DevicePlan *p, double *kernel_ms, unsigned long long *term_apps, std::string &err) {
    DeviceCtx *c = p->ctx;
    HIP_TRY(hipSetDevice(c->device));
    if (p->n_out == 0) {
        if (kernel_ms) *kernel_ms = 0.0;
        if (term_apps) *term_apps = 0;
        return 0;
    }
    // counters cleared by a kernel (not a memset node, so the launch sequence
    // can be captured into a graph)
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, p->args.counters, 8);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    int rc = 0;
    switch (p->kp1) {
        case 2: rc = launch_n<2>(p->args, p->n_out, p->wmax, c->stream, err); break;
        case 3: rc = launch_n<3>(p->args, p->n_out, p->wmax, c->stream, err); break;
        case 4: rc = launch_n<4>(p->args, p->n_out, p->wmax, c->stream, err); break;
        case 5: rc = launch_n<5>(p->args, p->n_out, p->wmax, c->stream, err); break;
        case 6: rc = launch_n<6>(p->args, p->n_out, p->wmax, c->stream, err); break;
        case 7: rc = launch_n<7>(p->args, p->n_out, p->wmax, c->stream, err); break;
        case 8: rc = launch_n<8>(p->args, p->n_out, p->wmax, c->stream, err); break;
        case 9: rc = launch_n<9>(p->args, p->n_out, p->wmax, c->stream, err); break;
        case 10: rc = launch_n<10>(p->args, p->n_out, p->wmax, c->stream, err); break;
        case 11: rc = launch_n<11>(p->args, p->n_out, p->wmax, c->stream, err); break;
        case 12: rc = launch_n<12>(p->args, p->n_out, p->wmax, c->stream, err); break;
        default: err = "unsupported basis size"; return -1;
    }
    if (rc) return rc;
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    unsigned long long cnt[8];
    HIP_TRY(hipMemcpyAsync(cnt, p->args.counters, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (kernel_ms) *kernel_ms = ms;
    if (term_apps) *term_apps = cnt[0];
    return 0;
}

int dev_plan_download(DevicePlan *p, double *out_ns, std::string &err) {
    DeviceCtx *c = p->ctx;
    HIP_TRY(hipSetDevice(c->device));
    if (p->n_out == 0) return 0;
    HIP_TRY(hipMemcpyAsync(out_ns, p->args.out_ns, sizeof(double) * (size_t)p->n_out * 8 * p->kp1 * p->kp1 * p->kp1,
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int dev_apply_plan(DeviceCtx *c, const void *op_key, const OperFlat &f, const ApplyPlan &plan, double *out_ns,
                   double *kernel_ms, unsigned long long *term_apps, std::string &err) {
    DevicePlan *p = nullptr;
    int rc = dev_plan_create(c, op_key, f, plan, &p, err);
    if (rc) return rc;
    rc = dev_plan_run(p, kernel_ms, term_apps, err);
    if (!rc) rc = dev_plan_download(p, out_ns, err);
    dev_plan_destroy(p);
    return rc;
}

}  // namespace mwg
