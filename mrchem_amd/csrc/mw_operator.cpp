// Non-standard-form operator tables for separated Gaussian kernels.
//
// Upstream MRCPP keeps one 1D OperatorTree per separated term
// (ConvolutionOperator / OperatorTree / BandWidth, fetched, absent here).  The
// tables here hold the same information flat: for every term m and depth d the
// four 1D NS components (s<-s, s<-d, d<-s, d<-d) for each translation l, as
// (k+1)x(k+1) blocks.  The s<-s blocks T^d_l are computed directly from the
// cross-correlation of the scaling functions against the Gaussian,
//     T^d_l(i,j) = h c int_{-1}^{1} psi_ij(u) exp(-alpha h^2 (u-l)^2) du,
// and the NS components at depth d from T^{d+1} through the two-scale filter:
//     [[ss, sd], [ds, dd]]^d_l = F [[T_{2l}, T_{2l+1}], [T_{2l-1}, T_{2l}]]^{d+1} F^T.
#include "mw_internal.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace mwg {

void t_block(const MRA &mra, double alpha, double coef1d, int depth, long long l, double *out) {
    const Basis &b = mra.basis;
    const int n = b.kp1, n2 = 2 * n;
    double h = mra.box_size(depth);
    double A = alpha * h * h;
    double pref = h * coef1d;
    std::vector<double> M1(n2), M2(n2);
    gauss_moments01(n2, A, (double)l, M1.data());
    gauss_moments01(n2, A, (double)(l + 1), M2.data());
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            const double *cij = &b.C[(i * n + j) * n2];
            const double *cji = &b.C[(j * n + i) * n2];
            double s = 0.0;
            for (int p = 0; p < n2; p++) s += cij[p] * M1[p] + ((p & 1) ? -cji[p] : cji[p]) * M2[p];
            out[i * n + j] = pref * s;
        }
    }
}

namespace {

double frob(const double *a, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * a[i];
    return std::sqrt(s);
}

// Lazily computed T^{d+1}_Delta blocks, Delta >= 0; negative via transpose.
struct TCache {
    const MRA *mra;
    double alpha, coef1d;
    int depth;
    std::vector<std::vector<double>> blk;
    const std::vector<double> &get(long long delta) {
        size_t idx = (size_t)delta;
        if (blk.size() <= idx) blk.resize(idx + 1);
        if (blk[idx].empty()) {
            int n = mra->kp1();
            blk[idx].resize(n * n);
            t_block(*mra, alpha, coef1d, depth, delta, blk[idx].data());
        }
        return blk[idx];
    }
    void fetch(long long delta, double *out) {
        int n = mra->kp1();
        if (delta >= 0) {
            const auto &b = get(delta);
            std::copy(b.begin(), b.end(), out);
        } else {
            const auto &b = get(-delta);
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) out[i * n + j] = b[j * n + i];
        }
    }
};

// NS components for translation l at depth d from T^{d+1}: out[c][i][j].
void ns_blocks(const MRA &mra, TCache &tc, long long l, double *out) {
    const int n = mra.kp1(), n2 = 2 * n;
    const std::vector<double> &F = mra.basis.F;
    std::vector<double> Mf(n2 * n2), tmp(n2 * n2), blk(n * n);
    auto place = [&](long long delta, int r0, int c0) {
        tc.fetch(delta, blk.data());
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) Mf[(r0 + i) * n2 + c0 + j] = blk[i * n + j];
    };
    place(2 * l, 0, 0);
    place(2 * l + 1, 0, n);
    place(2 * l - 1, n, 0);
    place(2 * l, n, n);
    // tmp = F Mf ; R = tmp F^T
    for (int i = 0; i < n2; i++)
        for (int j = 0; j < n2; j++) {
            double s = 0.0;
            for (int q = 0; q < n2; q++) s += F[i * n2 + q] * Mf[q * n2 + j];
            tmp[i * n2 + j] = s;
        }
    for (int i = 0; i < n2; i++)
        for (int j = 0; j < n2; j++) {
            double s = 0.0;
            for (int q = 0; q < n2; q++) s += tmp[i * n2 + q] * F[j * n2 + q];
            Mf[i * n2 + j] = s;
        }
    // c = 2*gt + ft : row half = gt (out s/d), column half = ft (in s/d)
    for (int c = 0; c < 4; c++) {
        int r0 = (c >> 1) * n, c0 = (c & 1) * n;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) out[c * n * n + i * n + j] = Mf[(r0 + i) * n2 + c0 + j];
    }
}

// Storage floors: a block is kept when it can matter at relative level tau.
// Wavelet components (c = 1, 2, 3) are kept above tau * ref; the s<-s
// component enters the 3D product only next to a wavelet component in another
// direction (except at the root), so off the root it is kept above
// tau * ref^2 / wmax, with ref the largest block norm and wmax the largest
// wavelet-component norm of the term at this depth.
void build_term_depth(const MRA &mra, OperTerm &term, int d, double tau) {
    const int n = mra.kp1(), nn = n * n;
    OperDepth &od = term.depth[d];
    od.L = -1;
    od.blocks.clear();
    od.norms.clear();
    long long Lscan = 0;
    double peak = 0.0;
    if (!term_depth_scan(mra, term.alpha, d, tau, Lscan, peak)) return;
    TCache tc{&mra, term.alpha, term.coef1d, d + 1, {}};
    std::vector<std::vector<double>> pos, neg;  // [l] -> 4*nn
    std::vector<double> pn, nnrm;               // norms [l][c]
    double ref = 0.0, wmax = 0.0;
    int quiet = 0;
    for (long long l = 0; l <= Lscan; l++) {
        std::vector<double> bp(4 * nn), bn(4 * nn);
        ns_blocks(mra, tc, l, bp.data());
        ns_blocks(mra, tc, -l, bn.data());
        double wrow = 0.0, srow = 0.0;
        for (int c = 0; c < 4; c++) {
            double a = frob(&bp[c * nn], nn), b = frob(&bn[c * nn], nn);
            pn.push_back(a);
            nnrm.push_back(b);
            ref = std::max(ref, std::max(a, b));
            if (c == 0) srow = std::max(a, b);
            else wrow = std::max(wrow, std::max(a, b));
        }
        wmax = std::max(wmax, std::max(wrow, 0.0));
        pos.push_back(std::move(bp));
        neg.push_back(std::move(bn));
        double fs = (d == 0) ? tau * ref : tau * ref * ref / std::max(wmax, 1e-300);
        if (wrow > tau * ref || srow > fs) quiet = 0;
        else quiet++;
        if (quiet >= 3 && (double)l >= peak) break;
    }
    double floor_w = tau * ref;
    double floor_s = (d == 0) ? tau * ref : tau * ref * ref / std::max(wmax, 1e-300);
    auto keep = [&](int c, double nr) { return nr > (c == 0 ? floor_s : floor_w); };
    long long L = -1;
    for (long long l = 0; l < (long long)pos.size(); l++)
        for (int c = 0; c < 4; c++)
            if (keep(c, pn[l * 4 + c]) || keep(c, nnrm[l * 4 + c])) L = std::max(L, l);
    if (L < 0) return;
    od.L = (int)L;
    int nl = 2 * (int)L + 1;
    od.blocks.assign((size_t)4 * nl * nn, 0.0);
    od.norms.assign((size_t)4 * nl, 0.0);
    for (int c = 0; c < 4; c++) {
        for (long long l = -L; l <= L; l++) {
            const double *src = (l >= 0) ? &pos[l][c * nn] : &neg[-l][c * nn];
            double nr = (l >= 0) ? pn[l * 4 + c] : nnrm[-l * 4 + c];
            if (!keep(c, nr)) continue;
            size_t idx = (size_t)c * nl + (size_t)(l + L);
            std::copy(src, src + nn, od.blocks.begin() + idx * nn);
            od.norms[idx] = nr;
        }
    }
}

}  // namespace

bool term_depth_scan(const MRA &mra, double alpha, int d, double tau, long long &Lscan, double &peak) {
    const int n = mra.kp1();
    double hd = mra.box_size(d), hd1 = mra.box_size(d + 1);
    double Ad = alpha * hd * hd, A1 = alpha * hd1 * hd1;
    if (d > 0) {
        // Wavelet components of a wide Gaussian are bounded by
        // A^{(k+1)/2} sqrt(2^{k+1}/(k+1)!) relative to its s<-s block.
        double lg = 0.5 * (n)*std::log(std::max(Ad, 1e-300)) + 0.5 * (n * std::log(2.0) - std::lgamma(n + 1.0));
        if (lg < std::log(0.1 * tau)) return false;
    }
    int bmax = std::max(mra.boxes[0], std::max(mra.boxes[1], mra.boxes[2]));
    long long Lw = ((long long)bmax << d) - 1;
    long long Ltail = 1 + (long long)std::ceil(std::sqrt(46.0 / std::max(A1, 1e-300)) / 2.0);
    Lscan = std::min(Lw, std::min(Ltail, (long long)1 << 22));
    peak = std::sqrt(2.0 * (2 * n) / std::max(A1, 1e-300)) / 2.0 + 2.0;
    return true;
}

double operator_tau(double build_prec) { return std::min(1e-10, 1e-4 * build_prec); }

void prepare_operator(Operator &op, const MRA &mra, double mu, double build_prec, int max_depth) {
    if (!(build_prec > 0)) throw std::invalid_argument("operator build precision must be > 0");
    if (max_depth < 0 || max_depth > mra.max_depth) max_depth = mra.max_depth;
    op.mra = &mra;
    op.mu = mu;
    op.build_prec = build_prec;
    op.max_depth = max_depth;
    double diag = 0.0;
    for (int d = 0; d < 3; d++) {
        double w = mra.world_hi(d) - mra.world_lo(d);
        diag += w * w;
    }
    op.r_max = std::sqrt(diag);
    double k_prec = build_prec / 10.0;
    double h_min = mra.box_size(mra.max_depth);
    op.r_min = std::sqrt(k_prec * h_min);
    std::vector<GaussTerm> g = fit_kernel(mu, k_prec, op.r_min, op.r_max, &op.fit_err);
    op.terms.assign(g.size(), OperTerm());
    for (size_t m = 0; m < g.size(); m++) {
        op.terms[m].alpha = g[m].alpha;
        op.terms[m].beta = g[m].beta;
        op.terms[m].coef1d = std::cbrt(g[m].beta);
        op.terms[m].depth.assign(max_depth + 1, OperDepth());
    }
}

void flatten_operator(const Operator &op, OperFlat &f) {
    const int n = op.mra->kp1(), nn = n * n;
    f = OperFlat();
    f.M = op.size();
    f.ndepth = op.max_depth + 1;
    f.kp1 = n;
    f.L.assign((size_t)f.M * f.ndepth, -1);
    f.boff.assign((size_t)f.M * f.ndepth, 0);
    f.noff.assign((size_t)f.M * f.ndepth, 0);
    long long nb = 0;
    for (int m = 0; m < f.M; m++)
        for (int d = 0; d < f.ndepth; d++) {
            const OperDepth &od = op.terms[m].depth[d];
            size_t i = (size_t)m * f.ndepth + d;
            f.L[i] = od.L;
            f.boff[i] = nb * nn;
            f.noff[i] = nb;
            if (od.L >= 0) nb += 4LL * (2 * od.L + 1);
        }
    f.blocks.assign((size_t)nb * nn, 0.0);
    f.norms.assign((size_t)nb, 0.0);
    for (int m = 0; m < f.M; m++)
        for (int d = 0; d < f.ndepth; d++) {
            const OperDepth &od = op.terms[m].depth[d];
            if (od.L < 0) continue;
            size_t i = (size_t)m * f.ndepth + d;
            std::copy(od.blocks.begin(), od.blocks.end(), f.blocks.begin() + f.boff[i]);
            std::copy(od.norms.begin(), od.norms.end(), f.norms.begin() + f.noff[i]);
        }
}

void build_operator_terms(Operator &op, const MRA &mra, const std::vector<GaussTerm> &g, int max_depth, double tau) {
    if (max_depth < 0 || max_depth > mra.max_depth) max_depth = mra.max_depth;
    op.mra = &mra;
    op.max_depth = max_depth;
    op.terms.assign(g.size(), OperTerm());
    for (size_t m = 0; m < g.size(); m++) {
        op.terms[m].alpha = g[m].alpha;
        op.terms[m].beta = g[m].beta;
        op.terms[m].coef1d = std::cbrt(g[m].beta);
        op.terms[m].depth.assign(max_depth + 1, OperDepth());
    }
    int nm = (int)g.size();
    int total = nm * (max_depth + 1);
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 1)
    for (int t = 0; t < total; t++) {
        int m = t / (max_depth + 1), d = t % (max_depth + 1);
        build_term_depth(mra, op.terms[m], d, tau);
    }
}

void build_operator(Operator &op, const MRA &mra, double mu, double build_prec, int max_depth) {
    prepare_operator(op, mra, mu, build_prec, max_depth);
    std::vector<GaussTerm> g(op.terms.size());
    for (size_t m = 0; m < g.size(); m++) g[m] = {op.terms[m].alpha, op.terms[m].beta};
    build_operator_terms(op, mra, g, op.max_depth, operator_tau(build_prec));
}

}  // namespace mwg
