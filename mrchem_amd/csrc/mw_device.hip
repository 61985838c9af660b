// HIP side of the multiwavelet Green's-function apply (gfx950).
//
// The per-node calculator of upstream MRCPP's OperApplicationCalculator
// (called from mrcpp::apply, src/qmoperators/two_electron/CoulombPotential.cpp:131)
// runs here: one workgroup per output node g.  For every band node f the
// workgroup screens all (term, ft, gt) candidates in parallel with the FP64
// bound ||O_x|| ||O_y|| ||O_z|| ||f_ft|| > thr_g (same operation order as the
// host oracle, so the surviving set is identical), then applies the surviving
// term-applications as three passes of (k+1)x(k+1) contractions through LDS,
// sharing the x pass between the output components that agree in x and the y
// pass between those that agree in x and y.  Accumulation is FP64 in
// registers; each thread owns a fixed strided slice of the 8(k+1)^3 output.
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

constexpr int kThreads = 256;
constexpr int kMaxActive = 512;

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

template <int N>
__global__ __launch_bounds__(kThreads) void k_apply_fp64(KArgs a) {
    constexpr int N2 = N * N, N3 = N2 * N, NQ = (N3 + kThreads - 1) / kThreads;
    extern __shared__ double smem[];
    double *fblk = smem;              // N3
    double *X = fblk + N3;            // 2 * N3
    double *Y = X + 2 * N3;           // 4 * N3
    double *Ob = Y + 4 * N3;          // [dim][gbit][N2]
    __shared__ unsigned int mask_s[2 * kMaxActive];

    const int g = blockIdx.x, tid = threadIdx.x;
    const int d = a.out_depth[g];
    const double thr = a.out_thr[g];
    const int w = a.band_w[g], side = 2 * w + 1;
    const long long boff = a.band_off[g];
    const int a0 = a.act_off[d], nact = a.act_off[d + 1] - a0;
    const int *act = a.act_terms + a0;

    double acc[8][NQ];
#pragma unroll
    for (int t = 0; t < 8; t++)
#pragma unroll
        for (int q = 0; q < NQ; q++) acc[t][q] = 0.0;
    unsigned long long napps = 0;

    const int nband = side * side * side;
    for (int p = 0; p < nband; p++) {
        const int s = a.band[boff + p];
        if (s < 0) continue;
        const int lx = p % side - w, ly = (p / side) % side - w, lz = p / (side * side) - w;
        double fn[8];
#pragma unroll
        for (int t = 0; t < 8; t++) fn[t] = a.slot_norm[(long long)s * 8 + t];
        for (int i = tid; i < 2 * nact; i += kThreads) mask_s[i] = 0u;
        __syncthreads();
        for (int idx = tid; idx < nact * 64; idx += kThreads) {
            const int ai = idx >> 6, ft = (idx >> 3) & 7, gt = idx & 7;
            if (!(fn[ft] > 0.0)) continue;
            if (d > 0 && ft == 0 && gt == 0) continue;
            const int m = act[ai];
            const int cx = ((gt & 1) << 1) | (ft & 1);
            const int cy = (((gt >> 1) & 1) << 1) | ((ft >> 1) & 1);
            const int cz = (((gt >> 2) & 1) << 1) | ((ft >> 2) & 1);
            const int *bwm = a.bw + ((long long)d * a.M + m) * 4;
            if (abs(lx) > bwm[cx] || abs(ly) > bwm[cy] || abs(lz) > bwm[cz]) continue;
            const double nx = op_norm(a, m, d, cx, lx);
            const double ny = op_norm(a, m, d, cy, ly);
            const double nz = op_norm(a, m, d, cz, lz);
            const double bound = ((nx * ny) * nz) * fn[ft];
            if (bound > thr) {
                atomicOr(&mask_s[ai * 2 + (ft >> 2)], 1u << ((ft & 3) * 8 + gt));
                napps++;
            }
        }
        __syncthreads();
        const double *fsrc = a.pool + a.slot_off[s];
        for (int ft = 0; ft < 8; ft++) {
            if (!(fn[ft] > 0.0)) continue;
            bool loaded = false;
            for (int ai = 0; ai < nact; ai++) {
                const unsigned int msk = (mask_s[ai * 2 + (ft >> 2)] >> ((ft & 3) * 8)) & 0xffu;
                if (msk == 0u) continue;
                const int m = act[ai];
                if (!loaded) {
                    for (int e = tid; e < N3; e += kThreads) fblk[e] = fsrc[ft * N3 + e];
                    loaded = true;
                }
                // which g-bits are needed per direction
                unsigned int needx = 0, needy = 0, needz = 0;
#pragma unroll
                for (int gt = 0; gt < 8; gt++)
                    if ((msk >> gt) & 1u) {
                        needx |= 1u << (gt & 1);
                        needy |= 1u << ((gt >> 1) & 1);
                        needz |= 1u << ((gt >> 2) & 1);
                    }
                const int L = a.opL[m * a.ndepth + d];
                const double *ob = a.op_blocks + a.op_boff[m * a.ndepth + d];
                const int ll[3] = {lx, ly, lz};
                const unsigned int need[3] = {needx, needy, needz};
                for (int e = tid; e < 6 * N2; e += kThreads) {
                    const int dim = e / (2 * N2), gb = (e / N2) & 1, r = e % N2;
                    if (!((need[dim] >> gb) & 1u)) continue;
                    const int fb = (ft >> dim) & 1;
                    const int c = (gb << 1) | fb;
                    Ob[e] = ob[((long long)c * (2 * L + 1) + (ll[dim] + L)) * N2 + r];
                }
                __syncthreads();
                // x pass: X[gx][jz][jy][ix] = sum_jx Ox[gx][ix][jx] f[jz][jy][jx]
                for (int e = tid; e < 2 * N3; e += kThreads) {
                    const int gx = e / N3, r = e % N3;
                    if (!((needx >> gx) & 1u)) continue;
                    const int ix = r % N, row = r / N;
                    const double *o = Ob + (0 * 2 + gx) * N2 + ix * N;
                    const double *f = fblk + row * N;
                    double v = 0.0;
#pragma unroll
                    for (int j = 0; j < N; j++) v = fma(o[j], f[j], v);
                    X[e] = v;
                }
                __syncthreads();
                // y pass: Y[gx + 2 gy][jz][iy][ix] = sum_jy Oy[gy][iy][jy] X[gx][jz][jy][ix]
                for (int e = tid; e < 4 * N3; e += kThreads) {
                    const int v4 = e / N3, r = e % N3;
                    const int gx = v4 & 1, gy = v4 >> 1;
                    if (!((needx >> gx) & 1u) || !((needy >> gy) & 1u)) continue;
                    const int ix = r % N, iy = (r / N) % N, jz = r / N2;
                    const double *o = Ob + (1 * 2 + gy) * N2 + iy * N;
                    const double *x = X + gx * N3 + jz * N2 + ix;
                    double v = 0.0;
#pragma unroll
                    for (int j = 0; j < N; j++) v = fma(o[j], x[j * N], v);
                    Y[e] = v;
                }
                __syncthreads();
                // z pass into the register accumulators
#pragma unroll
                for (int gt = 0; gt < 8; gt++) {
                    if (!((msk >> gt) & 1u)) continue;
                    const double *o = Ob + (2 * 2 + ((gt >> 2) & 1)) * N2;
                    const double *y = Y + ((gt & 1) + 2 * ((gt >> 1) & 1)) * N3;
#pragma unroll
                    for (int q = 0; q < NQ; q++) {
                        const int e = tid + q * kThreads;
                        if (e < N3) {
                            const int iz = e / N2, r = e % N2;
                            double v = 0.0;
#pragma unroll
                            for (int j = 0; j < N; j++) v = fma(o[iz * N + j], y[j * N2 + r], v);
                            acc[gt][q] += v;
                        }
                    }
                }
                __syncthreads();
            }
            if (loaded) __syncthreads();
        }
    }
    // one atomic per wavefront for the work counter
    for (int off = 32; off > 0; off >>= 1) napps += __shfl_down(napps, off);
    if ((tid & 63) == 0 && napps) atomicAdd(&a.counters[0], napps);
    double *dst = a.out_ns + (long long)g * 8 * N3;
#pragma unroll
    for (int t = 0; t < 8; t++)
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int e = tid + q * kThreads;
            if (e < N3) dst[t * N3 + e] = acc[t][q];
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
    int n_out = 0, kp1 = 0;
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
static int launch_n(const KArgs &a, int n_out, hipStream_t s, std::string &err) {
    constexpr int N2 = N * N, N3 = N2 * N;
    size_t shmem = sizeof(double) * (7 * N3 + 6 * N2);
    HIP_TRY(hipFuncSetAttribute((const void *)k_apply_fp64<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    hipLaunchKernelGGL(k_apply_fp64<N>, dim3(n_out), dim3(kThreads), shmem, s, a);
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
        case 2: rc = launch_n<2>(p->args, p->n_out, c->stream, err); break;
        case 3: rc = launch_n<3>(p->args, p->n_out, c->stream, err); break;
        case 4: rc = launch_n<4>(p->args, p->n_out, c->stream, err); break;
        case 5: rc = launch_n<5>(p->args, p->n_out, c->stream, err); break;
        case 6: rc = launch_n<6>(p->args, p->n_out, c->stream, err); break;
        case 7: rc = launch_n<7>(p->args, p->n_out, c->stream, err); break;
        case 8: rc = launch_n<8>(p->args, p->n_out, c->stream, err); break;
        case 9: rc = launch_n<9>(p->args, p->n_out, c->stream, err); break;
        case 10: rc = launch_n<10>(p->args, p->n_out, c->stream, err); break;
        case 11: rc = launch_n<11>(p->args, p->n_out, c->stream, err); break;
        case 12: rc = launch_n<12>(p->args, p->n_out, c->stream, err); break;
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
