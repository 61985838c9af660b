// Host FunctionTree: Morton-keyed 2^3-tree of compressed 8(k+1)^3 blocks,
// two-scale transforms, norms, dot products, Gaussian projection and the
// wavelet split rule.
//
// Node layout (8 components of (k+1)^3, x fastest, component bit d set when
// direction d is a wavelet) matches the MRCPP node block MRChem reads in
// src/mrdft/Functional.cpp:215 (ncoefs = tDim * kp1^D).  Every node stores its
// full compressed [s, d] block, as upstream MRCPP nodes do.
#include "mw_internal.hpp"

#include <sys/mman.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <new>
#include <stdexcept>

namespace mwg {

void *big_alloc(size_t bytes) {
    if (bytes < kBigAllocMin) return ::operator new(bytes);
    const size_t sz = (bytes + kHugePage - 1) / kHugePage * kHugePage;
    void *p = std::aligned_alloc(kHugePage, sz);
    if (!p) throw std::bad_alloc();
    (void)madvise(p, sz, MADV_HUGEPAGE);  // a hint: without THP the pages stay 4 KiB
    return p;
}

void big_free(void *p, size_t bytes) {
    if (!p) return;
    if (bytes < kBigAllocMin)
        ::operator delete(p);
    else
        std::free(p);
}

uint64_t morton3(uint64_t x, uint64_t y, uint64_t z) {
    auto spread = [](uint64_t v) {
        v &= 0x1fffffULL;
        v = (v | (v << 32)) & 0x1f00000000ffffULL;
        v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
        v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
        v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
        v = (v | (v << 2)) & 0x1249249249249249ULL;
        return v;
    };
    return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

uint64_t Tree::morton(const NodeKey &k) const {
    uint64_t o[3];
    for (int d = 0; d < 3; d++) o[d] = (uint64_t)(k.l[d] - mra->tlo(k.depth, d));
    return morton3(o[0], o[1], o[2]);
}

int Tree::find(int depth, const long long *l) const {
    if (depth < 0 || !mra->in_world(depth, l)) return -1;
    NodeKey k{depth, {l[0], l[1], l[2]}};
    auto it = index.find({depth, morton(k)});
    return it == index.end() ? -1 : it->second;
}

int Tree::add_node(const NodeKey &k, int par) {
    if (k.depth > 20) throw std::runtime_error("tree depth beyond 20 is not supported");
    int id = (int)keys.size();
    keys.push_back(k);
    parent.push_back(par);
    child0.push_back(-1);
    if (!structure_only) coefs.resize(coefs.size() + mra->block(), 0.0);
    index[{k.depth, morton(k)}] = id;
    return id;
}

void Tree::clear() {
    keys.clear();
    parent.clear();
    child0.clear();
    coefs.clear();
    index.clear();
    sqnorm = 0.0;
}

void Tree::make_roots() {
    for (long long z = mra->corner[2]; z < mra->corner[2] + mra->boxes[2]; z++)
        for (long long y = mra->corner[1]; y < mra->corner[1] + mra->boxes[1]; y++)
            for (long long x = mra->corner[0]; x < mra->corner[0] + mra->boxes[0]; x++)
                add_node(NodeKey{0, {x, y, z}}, -1);
}

int Tree::split(int node) {
    if (child0[node] >= 0) return child0[node];
    NodeKey k = keys[node];
    int first = -1;
    for (int t = 0; t < 8; t++) {
        NodeKey c{k.depth + 1, {2 * k.l[0] + (t & 1), 2 * k.l[1] + ((t >> 1) & 1), 2 * k.l[2] + ((t >> 2) & 1)}};
        int id = add_node(c, node);
        if (t == 0) first = id;
    }
    child0[node] = first;
    return first;
}

int Tree::max_depth() const {
    int m = 0;
    for (auto &k : keys) m = std::max(m, k.depth);
    return m;
}

int Tree::n_leaves() const {
    int n = 0;
    for (int c : child0) n += (c < 0);
    return n;
}

// ----------------------------------------------------------- transforms
namespace {
// Apply the 2kp1 x 2kp1 matrix (or its transpose) along direction dim to the
// pair of blocks (lo, hi).
void pair_transform(const Basis &b, double *lo, double *hi, int dim, bool transpose) {
    const int n = b.kp1, n2 = 2 * n;
    const double *F = b.F.data();
    int stride = (dim == 0) ? 1 : (dim == 1 ? n : n * n);
    double v[64], u[64];
    for (int a = 0; a < n; a++) {
        for (int c = 0; c < n; c++) {
            // base offset of the line: the two other indices are (a, c)
            int base;
            if (dim == 0) base = a * n + c * n * n;
            else if (dim == 1) base = a + c * n * n;
            else base = a + c * n;
            for (int i = 0; i < n; i++) {
                v[i] = lo[base + i * stride];
                v[n + i] = hi[base + i * stride];
            }
            for (int i = 0; i < n2; i++) {
                double s = 0.0;
                if (!transpose)
                    for (int j = 0; j < n2; j++) s += F[i * n2 + j] * v[j];
                else
                    for (int j = 0; j < n2; j++) s += F[j * n2 + i] * v[j];
                u[i] = s;
            }
            for (int i = 0; i < n; i++) {
                lo[base + i * stride] = u[i];
                hi[base + i * stride] = u[n + i];
            }
        }
    }
}
}  // namespace

void mw_compress(const Basis &b, const double *in, double *out) {
    const int n3 = b.kp1 * b.kp1 * b.kp1;
    if (out != in) std::copy(in, in + 8 * n3, out);
    for (int dim = 0; dim < 3; dim++)
        for (int t = 0; t < 8; t++)
            if (!((t >> dim) & 1)) pair_transform(b, out + t * n3, out + (t | (1 << dim)) * n3, dim, false);
}

void mw_reconstruct(const Basis &b, const double *in, double *out) {
    const int n3 = b.kp1 * b.kp1 * b.kp1;
    if (out != in) std::copy(in, in + 8 * n3, out);
    for (int dim = 2; dim >= 0; dim--)
        for (int t = 0; t < 8; t++)
            if (!((t >> dim) & 1)) pair_transform(b, out + t * n3, out + (t | (1 << dim)) * n3, dim, true);
}

namespace {
// One direction of the inverse two-scale step towards child half `bit`:
// out[i] = sum_j H_bit[j][i] lo[j] + G_bit[j][i] hi[j] along dim, where
// H_bit, G_bit are the (bit) column halves of the filter rows [H0 H1], [G0 G1].
// hi == nullptr drops the wavelet half.
void child_step(const Basis &b, const double *lo, const double *hi, int dim, int bit, double *out) {
    const int n = b.kp1, n2 = 2 * n, nn = n * n;
    const double *F = b.F.data();
    // Each output element sums j = 0..n-1 over lo, then over hi, in this
    // order in every branch; the z and y branches run it for many elements
    // at once along contiguous memory (out must not alias lo or hi).
    if (dim == 2) {
        for (int i = 0; i < n; i++) {
            double *o = out + (size_t)i * nn;
            for (int e = 0; e < nn; e++) o[e] = 0.0;
            for (int j = 0; j < n; j++) {
                const double f = F[j * n2 + bit * n + i], *l = lo + (size_t)j * nn;
                for (int e = 0; e < nn; e++) o[e] += f * l[e];
            }
            if (hi)
                for (int j = 0; j < n; j++) {
                    const double f = F[(n + j) * n2 + bit * n + i], *h = hi + (size_t)j * nn;
                    for (int e = 0; e < nn; e++) o[e] += f * h[e];
                }
        }
        return;
    }
    if (dim == 1) {
        for (int c = 0; c < n; c++) {
            const double *L = lo + (size_t)c * nn, *H = hi ? hi + (size_t)c * nn : nullptr;
            for (int i = 0; i < n; i++) {
                double *o = out + (size_t)c * nn + i * n;
                for (int x = 0; x < n; x++) o[x] = 0.0;
                for (int j = 0; j < n; j++) {
                    const double f = F[j * n2 + bit * n + i];
                    for (int x = 0; x < n; x++) o[x] += f * L[j * n + x];
                }
                if (H)
                    for (int j = 0; j < n; j++) {
                        const double f = F[(n + j) * n2 + bit * n + i];
                        for (int x = 0; x < n; x++) o[x] += f * H[j * n + x];
                    }
            }
        }
        return;
    }
    // x: lines of n contiguous values; the i loop runs over a row of F
    for (int r = 0; r < nn; r++) {
        const double *v = lo + (size_t)r * n, *u = hi ? hi + (size_t)r * n : nullptr;
        double *o = out + (size_t)r * n;
        for (int i = 0; i < n; i++) o[i] = 0.0;
        for (int j = 0; j < n; j++) {
            const double *f = F + j * n2 + bit * n;
            const double vj = v[j];
            for (int i = 0; i < n; i++) o[i] += f[i] * vj;
        }
        if (u)
            for (int j = 0; j < n; j++) {
                const double *f = F + (n + j) * n2 + bit * n;
                const double uj = u[j];
                for (int i = 0; i < n; i++) o[i] += f[i] * uj;
            }
    }
}
}  // namespace

void mw_child_scaling(const Basis &b, const double *in, bool s_only, int t, double *out) {
    const int n3 = b.kp1 * b.kp1 * b.kp1;
    if (s_only) {
        std::vector<double> t1(n3), t2(n3);
        child_step(b, in, nullptr, 2, (t >> 2) & 1, t1.data());
        child_step(b, t1.data(), nullptr, 1, (t >> 1) & 1, t2.data());
        child_step(b, t2.data(), nullptr, 0, t & 1, out);
        return;
    }
    // z halves: components c and c|4 -> 4 blocks keyed by the (x, y) bits of c
    std::vector<double> z(4 * (size_t)n3), y(2 * (size_t)n3);
    for (int c = 0; c < 4; c++) child_step(b, in + c * n3, in + (c | 4) * n3, 2, (t >> 2) & 1, z.data() + c * n3);
    for (int c = 0; c < 2; c++) child_step(b, z.data() + c * n3, z.data() + (c | 2) * n3, 1, (t >> 1) & 1, y.data() + c * n3);
    child_step(b, y.data(), y.data() + n3, 0, t & 1, out);
}

void component_norms(const MRA &mra, const double *blk, double *norms8) {
    const int n3 = mra.kp1_3();
    for (int t = 0; t < 8; t++) {
        double s = 0.0;
        for (int i = 0; i < n3; i++) s += blk[t * n3 + i] * blk[t * n3 + i];
        norms8[t] = std::sqrt(s);
    }
}

double tree_square_norm(const Tree &t) {
    const int n3 = t.mra->kp1_3();
    double s = 0.0;
    for (int i = 0; i < t.size(); i++) {
        const double *b = t.node(i);
        int start = (t.parent[i] < 0) ? 0 : n3;
        for (int j = start; j < 8 * n3; j++) s += b[j] * b[j];
    }
    return s;
}

double tree_dot(const Tree &a, const Tree &b) {
    const int n3 = a.mra->kp1_3();
    double s = 0.0;
    for (int i = 0; i < a.size(); i++) {
        int j = b.find(a.keys[i].depth, a.keys[i].l);
        if (j < 0) continue;
        const double *x = a.node(i), *y = b.node(j);
        int start = (a.parent[i] < 0) ? 0 : n3;
        for (int q = start; q < 8 * n3; q++) s += x[q] * y[q];
    }
    return s;
}

void tree_compress(Tree &t) {
    const int n3 = t.mra->kp1_3(), blk = t.mra->block();
    int maxd = t.max_depth();
    std::vector<std::vector<int>> by_depth(maxd + 1);
    for (int i = 0; i < t.size(); i++)
        if (t.child0[i] >= 0) by_depth[t.keys[i].depth].push_back(i);
    for (int d = maxd; d >= 0; d--) {
        auto &v = by_depth[d];
#pragma omp parallel for num_threads(team_size()) schedule(static)
        for (int q = 0; q < (int)v.size(); q++) {
            int i = v[q];
            std::vector<double> in(blk);
            for (int c = 0; c < 8; c++) {
                const double *src = t.node(t.child0[i] + c);
                std::copy(src, src + n3, in.begin() + c * n3);
            }
            mw_compress(t.mra->basis, in.data(), t.node(i));
        }
    }
}

// ----------------------------------------------------------- Gaussians
double GaussSum::eval(const double *r) const {
    double s = 0.0;
    for (int i = 0; i < size(); i++) {
        double d2 = 0.0;
        for (int d = 0; d < 3; d++) {
            double x = r[d] - pos[3 * i + d];
            d2 += x * x;
        }
        s += coef[i] * std::pow(exp[i] / M_PI, 1.5) * std::exp(-exp[i] * d2);
    }
    return s;
}

double GaussSum::square_norm() const {
    double s = 0.0;
    for (int i = 0; i < size(); i++)
        for (int j = 0; j < size(); j++) {
            double a = exp[i], b = exp[j];
            double R2 = 0.0;
            for (int d = 0; d < 3; d++) {
                double x = pos[3 * i + d] - pos[3 * j + d];
                R2 += x * x;
            }
            double p = a * b / (a + b);
            s += coef[i] * coef[j] * std::pow(a * b / (M_PI * M_PI), 1.5) * std::pow(M_PI / (a + b), 1.5) *
                 std::exp(-p * R2);
        }
    return s;
}

namespace {
// 1D integrals I[i] = sqrt(h) sum_p L2B[i][p] int_0^1 P~_p(v) exp(-A (v - c)^2) dv
// of Gaussian g along direction d on translation l at depth; false when all
// vanish.
bool project_1d(const MRA &mra, const GaussSum &f, int depth, int d, long long l, int g, double *I) {
    const int n = mra.kp1();
    const double h = mra.box_size(depth), sh = std::sqrt(h);
    const double a = f.exp[g], A = a * h * h;
    const double c = f.pos[3 * g + d] / h - (double)l;
    double mom[32];
    gauss_moments01(n, A, c, mom);
    bool any = false;
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int p = 0; p < n; p++) s += mra.basis.L2B[i * n + p] * mom[p];
        I[i] = sh * s;
        any |= (s != 0.0);
    }
    return any;
}

// out[(z n + y) n + x] = sum_g c_g (a_g/pi)^{3/2} Iz[z] Iy[y] Ix[x]; Ig[d] points
// at the n integrals of Gaussian g along d, or is null when they all vanish.
void combine_1d(const MRA &mra, const GaussSum &f, const double *const *Ig, double *out) {
    const int n = mra.kp1(), n3 = n * n * n;
    std::fill(out, out + n3, 0.0);
    for (int g = 0; g < f.size(); g++) {
        const double *Ix = Ig[3 * g], *Iy = Ig[3 * g + 1], *Iz = Ig[3 * g + 2];
        if (!Ix || !Iy || !Iz) continue;
        const double pre = f.coef[g] * std::pow(f.exp[g] / M_PI, 1.5);
        for (int z = 0; z < n; z++)
            for (int y = 0; y < n; y++) {
                const double yz = pre * Iz[z] * Iy[y];
                for (int x = 0; x < n; x++) out[(z * n + y) * n + x] += yz * Ix[x];
            }
    }
}
}  // namespace

void project_scaling(const MRA &mra, const GaussSum &f, int depth, const long long *l, double *out) {
    const int n = mra.kp1();
    std::vector<double> I((size_t)f.size() * 3 * n);
    std::vector<const double *> Ig((size_t)f.size() * 3, nullptr);
    for (int g = 0; g < f.size(); g++)
        for (int d = 0; d < 3; d++) {
            double *dst = I.data() + ((size_t)g * 3 + d) * n;
            if (!project_1d(mra, f, depth, d, l[d], g, dst)) break;
            Ig[3 * g + d] = dst;
        }
    combine_1d(mra, f, Ig.data(), out);
}

bool split_check(const MRA &mra, int depth, double wnorm, double prec, double tnorm) {
    if (prec <= 0.0) return false;
    double scale = mra.n0 + depth;
    double thr = std::max(2.0 * 2.220446049250313e-16, prec * tnorm * std::pow(2.0, -0.5 * (scale + 1)));
    return wnorm > thr;
}

void project_adaptive(Tree &t, const GaussSum &f, double prec, bool abs_prec, int max_depth) {
    const MRA &mra = *t.mra;
    const int n = mra.kp1(), n3 = mra.kp1_3(), blk = mra.block(), ng = f.size();
    if (max_depth < 0 || max_depth > mra.max_depth) max_depth = mra.max_depth;
    t.clear();
    t.make_roots();
    double tnorm = abs_prec ? 1.0 : std::sqrt(std::max(f.square_norm(), 0.0));
    std::vector<int> work;
    for (int i = 0; i < t.size(); i++) work.push_back(i);
    while (!work.empty()) {
        // the nodes of one level share their depth; their children's 1D
        // integrals are tabulated once per (direction, translation, Gaussian)
        const int depth = t.keys[work[0]].depth, cd = depth + 1;
        std::vector<long long> lo(3), span(3);
        for (int d = 0; d < 3; d++) {
            long long a = 2 * t.keys[work[0]].l[d], b = a + 1;
            for (int i : work) {
                a = std::min(a, 2 * t.keys[i].l[d]);
                b = std::max(b, 2 * t.keys[i].l[d] + 1);
            }
            lo[d] = a;
            span[d] = b - a + 1;
        }
        const long long ntab = (span[0] + span[1] + span[2]) * ng;
        std::vector<double> tab((size_t)ntab * n);
        std::vector<char> nz((size_t)ntab, 0);
        auto tab_index = [&](int d, long long l, int g) -> long long {
            long long base = 0;
            for (int e = 0; e < d; e++) base += span[e];
            return (base + (l - lo[d])) * ng + g;
        };
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 16)
        for (long long e = 0; e < ntab; e++) {
            const int g = (int)(e % ng);
            long long r = e / ng;
            int d = 0;
            while (r >= span[d]) r -= span[d++];
            nz[e] = project_1d(mra, f, cd, d, lo[d] + r, g, tab.data() + e * n) ? 1 : 0;
        }
        std::vector<char> do_split(work.size(), 0);
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 4)
        for (int q = 0; q < (int)work.size(); q++) {
            int i = work[q];
            const NodeKey k = t.keys[i];
            std::vector<double> ch(blk);
            std::vector<const double *> Ig((size_t)ng * 3);
            for (int c = 0; c < 8; c++) {
                long long cl[3] = {2 * k.l[0] + (c & 1), 2 * k.l[1] + ((c >> 1) & 1), 2 * k.l[2] + ((c >> 2) & 1)};
                for (int g = 0; g < ng; g++) {
                    bool dead = false;
                    for (int d = 0; d < 3; d++) {
                        const long long e = tab_index(d, cl[d], g);
                        Ig[3 * g + d] = (!dead && nz[e]) ? tab.data() + e * n : nullptr;
                        dead = dead || !nz[e];
                    }
                }
                combine_1d(mra, f, Ig.data(), ch.data() + c * n3);
            }
            double *dst = t.coefs.data() + (size_t)i * blk;
            mw_compress(mra.basis, ch.data(), dst);
            double w = 0.0;
            for (int e = n3; e < blk; e++) w += dst[e] * dst[e];
            if (k.depth < max_depth && split_check(mra, k.depth, std::sqrt(w), prec, tnorm)) do_split[q] = 1;
        }
        std::vector<int> next;
        for (size_t q = 0; q < work.size(); q++) {
            if (!do_split[q]) continue;
            int c0 = t.split(work[q]);
            for (int c = 0; c < 8; c++) next.push_back(c0 + c);
        }
        work.swap(next);
    }
    t.sqnorm = tree_square_norm(t);
}

double tree_eval(const Tree &t, const double *r) {
    const MRA &mra = *t.mra;
    const int n = mra.kp1(), n3 = n * n * n;
    // locate the root box
    long long l[3];
    for (int d = 0; d < 3; d++) {
        double h = mra.box_size(0);
        l[d] = (long long)std::floor(r[d] / h);
        l[d] = std::max(mra.tlo(0, d), std::min(mra.thi(0, d) - 1, l[d]));
    }
    int node = t.find(0, l);
    if (node < 0) return 0.0;
    while (t.child0[node] >= 0) {
        int depth = t.keys[node].depth + 1;
        double h = mra.box_size(depth);
        int c = 0;
        for (int d = 0; d < 3; d++) {
            long long cl = (long long)std::floor(r[d] / h);
            cl = std::max(2 * t.keys[node].l[d], std::min(2 * t.keys[node].l[d] + 1, cl));
            c |= (int)(cl - 2 * t.keys[node].l[d]) << d;
        }
        node = t.child0[node] + c;
    }
    // reconstruct the leaf to its children's scaling coefficients
    std::vector<double> rec(8 * n3);
    mw_reconstruct(mra.basis, t.node(node), rec.data());
    int depth = t.keys[node].depth + 1;
    double h = mra.box_size(depth);
    int c = 0;
    double xi[3];
    for (int d = 0; d < 3; d++) {
        long long cl = (long long)std::floor(r[d] / h);
        cl = std::max(2 * t.keys[node].l[d], std::min(2 * t.keys[node].l[d] + 1, cl));
        c |= (int)(cl - 2 * t.keys[node].l[d]) << d;
        xi[d] = r[d] / h - (double)cl;
    }
    std::vector<double> px(n), py(n), pz(n);
    mra.basis.eval(std::min(1.0, std::max(0.0, xi[0])), px.data());
    mra.basis.eval(std::min(1.0, std::max(0.0, xi[1])), py.data());
    mra.basis.eval(std::min(1.0, std::max(0.0, xi[2])), pz.data());
    const double *s = rec.data() + c * n3;
    double v = 0.0;
    for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) v += s[(z * n + y) * n + x] * px[x] * py[y] * pz[z];
    return v * std::pow(h, -1.5);
}

}  // namespace mwg
