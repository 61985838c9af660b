// Quadrature, scaling bases, two-scale filters, Gaussian moments and
// cross-correlation coefficients.
//
// Basis definitions follow tools/legendre_to_interpolating_basis/
// legendre_to_interpolating.py:17-32 (Legendre scaling functions
// sqrt(2j+1) P_j(2x-1) and the interpolating functions built from them at the
// Gauss-Legendre roots).  The filters are the Alpert two-scale relations of the
// upstream MRCPP core (MWFilter/CrossCorrelation), computed here by exact
// quadrature instead of read from filter files.
#include "mw_internal.hpp"

#include <omp.h>
#include <sched.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>

namespace mwg {

namespace {
int cpus_usable() {
    int n = 0;
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
    // cgroup v2 quota "max" or "<quota> <period>"
    if (FILE *f = std::fopen("/sys/fs/cgroup/cpu.max", "r")) {
        char q[32] = {0};
        long long period = 0;
        if (std::fscanf(f, "%31s %lld", q, &period) == 2 && period > 0 && q[0] >= '0' && q[0] <= '9') {
            const long long quota = std::atoll(q);
            const int qc = (int)std::max<long long>(1, quota / period);
            n = n > 0 ? std::min(n, qc) : qc;
        }
        std::fclose(f);
    }
    return n;
}
}  // namespace

int team_size() {
    static const int n = [] {
        if (const char *s = std::getenv("MWG_NUM_THREADS")) {
            const int v = std::atoi(s);
            if (v > 0) return v;
        }
        const int omp = std::max(1, omp_get_max_threads());
        const int cpus = cpus_usable();
        return cpus > 0 ? std::max(1, std::min(omp, cpus / 2)) : omp;
    }();
    return n;
}

void gauss_legendre01(int n, double *x, double *w) {
    if (n < 1) throw std::invalid_argument("gauss_legendre01: n < 1");
    for (int i = 0; i < n; i++) {
        // Newton on P_n starting from the Tricomi estimate of the i-th root.
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int j = 1; j < n; j++) {
                double p2 = ((2.0 * j + 1.0) * z * p1 - j * p0) / (j + 1.0);
                p0 = p1;
                p1 = p2;
            }
            double pn = (n == 1) ? z : p1;
            double pm = (n == 1) ? 1.0 : p0;
            dp = n * (z * pn - pm) / (z * z - 1.0);
            double dz = pn / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-17) {
                // one more refinement for the derivative at the converged root
                p0 = 1.0;
                p1 = z;
                for (int j = 1; j < n; j++) {
                    double p2 = ((2.0 * j + 1.0) * z * p1 - j * p0) / (j + 1.0);
                    p0 = p1;
                    p1 = p2;
                }
                pn = (n == 1) ? z : p1;
                pm = (n == 1) ? 1.0 : p0;
                dp = n * (z * pn - pm) / (z * z - 1.0);
                break;
            }
        }
        // z runs from near +1 downwards; map to ascending nodes on [0,1]
        x[i] = 0.5 * (1.0 - z);
        w[i] = 1.0 / ((1.0 - z * z) * dp * dp);  // 2/((1-z^2)P'^2) halved for [0,1]
    }
}

void legendre01(int n, double x, double *out) {
    double t = 2.0 * x - 1.0;
    double p0 = 1.0, p1 = t;
    if (n > 0) out[0] = 1.0;
    if (n > 1) out[1] = std::sqrt(3.0) * t;
    for (int p = 1; p + 1 < n; p++) {
        double p2 = ((2.0 * p + 1.0) * t * p1 - p * p0) / (p + 1.0);
        p0 = p1;
        p1 = p2;
        out[p + 1] = std::sqrt(2.0 * (p + 1) + 1.0) * p2;
    }
}

namespace {
struct Rule {
    std::vector<double> x, w;
};
const Rule &cached_rule(int n) {
    static std::mutex mtx;
    static std::map<int, Rule> cache;
    std::lock_guard<std::mutex> g(mtx);
    auto it = cache.find(n);
    if (it != cache.end()) return it->second;
    Rule r;
    r.x.resize(n);
    r.w.resize(n);
    gauss_legendre01(n, r.x.data(), r.w.data());
    return cache.emplace(n, std::move(r)).first->second;
}
}  // namespace

void quad_rule01(int n, const double **x, const double **w) {
    const Rule &r = cached_rule(n);
    *x = r.x.data();
    *w = r.w.data();
}

void gauss_moments01(int n, double A, double c, double *out) {
    for (int p = 0; p < n; p++) out[p] = 0.0;
    if (A < 0) throw std::invalid_argument("gauss_moments01: A < 0");
    double a = 0.0, b = 1.0;
    if (A > 0.0) {
        // Keep the part of [0,1] where the Gaussian is within e^-46 of its
        // largest value on [0,1]; nothing below e^-745 is representable.
        double vstar = std::min(1.0, std::max(0.0, c));
        double dstar = std::fabs(vstar - c);
        if (A * dstar * dstar > 745.0) return;
        double R = std::sqrt(dstar * dstar + 46.0 / A);
        a = std::max(0.0, c - R);
        b = std::min(1.0, c + R);
        if (b <= a) return;
    }
    // Piece width so that the exponent varies by at most ~16 on each piece.
    double maxdist = std::max(std::fabs(a - c), std::fabs(b - c));
    double delta = b - a;
    if (A > 0.0) {
        delta = std::min(delta, 4.0 / std::sqrt(A));
        if (A * maxdist > 0.0) delta = std::min(delta, 8.0 / (A * maxdist));
    }
    int npieces = std::max(1, (int)std::ceil((b - a) / delta - 1e-12));
    if (npieces > 4096) npieces = 4096;
    double width = (b - a) / npieces;
    const Rule &r = cached_rule(n + 20);
    std::vector<double> P(n);
    for (int q = 0; q < npieces; q++) {
        double u0 = a + q * width;
        for (size_t i = 0; i < r.x.size(); i++) {
            double v = u0 + width * r.x[i];
            double g = std::exp(-A * (v - c) * (v - c)) * r.w[i] * width;
            if (g == 0.0) continue;
            legendre01(n, v, P.data());
            for (int p = 0; p < n; p++) out[p] += g * P[p];
        }
    }
}

void Basis::eval(double x, double *out) const {
    std::vector<double> P(kp1);
    legendre01(kp1, x, P.data());
    for (int i = 0; i < kp1; i++) {
        double s = 0.0;
        for (int p = 0; p < kp1; p++) s += L2B[i * kp1 + p] * P[p];
        out[i] = s;
    }
}

Basis make_basis(int k, int type) {
    if (k < 1 || k > 20) throw std::invalid_argument("basis order must be in [1,20]");
    if (type != BASIS_INTERPOLATING && type != BASIS_LEGENDRE)
        throw std::invalid_argument("unknown basis type");
    Basis b;
    b.k = k;
    b.kp1 = k + 1;
    b.type = type;
    const int n = b.kp1, n2 = 2 * n;
    b.L2B.assign(n * n, 0.0);
    if (type == BASIS_LEGENDRE) {
        for (int i = 0; i < n; i++) b.L2B[i * n + i] = 1.0;
    } else {
        // phi_j = sqrt(w_j) sum_i P~_i(x_j) P~_i : the Lagrange polynomial at
        // root x_j divided by sqrt(w_j).
        std::vector<double> x(n), w(n), P(n);
        gauss_legendre01(n, x.data(), w.data());
        for (int j = 0; j < n; j++) {
            legendre01(n, x[j], P.data());
            for (int i = 0; i < n; i++) b.L2B[j * n + i] = std::sqrt(w[j]) * P[i];
        }
    }
    // H0, H1 by exact quadrature (integrands have degree 2k).
    b.F.assign(n2 * n2, 0.0);
    {
        const Rule &r = cached_rule(n2);
        std::vector<double> pa(n), pb(n);
        for (size_t q = 0; q < r.x.size(); q++) {
            double t = r.x[q], wq = r.w[q] / std::sqrt(2.0);
            b.eval(t, pb.data());
            b.eval(0.5 * t, pa.data());
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) b.F[i * n2 + j] += wq * pa[i] * pb[j];
            b.eval(0.5 * (t + 1.0), pa.data());
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) b.F[i * n2 + n + j] += wq * pa[i] * pb[j];
        }
    }
    // G0, G1: orthonormal complement of the H rows.  Candidates are the
    // left-half fine scaling functions; each is orthogonalised twice against
    // the H rows and the previous wavelets (modified Gram-Schmidt).
    for (int g = 0; g < n; g++) {
        std::vector<double> v(n2, 0.0);
        v[g] = 1.0;
        for (int pass = 0; pass < 2; pass++) {
            for (int r = 0; r < n + g; r++) {
                const double *row = &b.F[r * n2];
                double dot = 0.0;
                for (int j = 0; j < n2; j++) dot += row[j] * v[j];
                for (int j = 0; j < n2; j++) v[j] -= dot * row[j];
            }
        }
        double nrm = 0.0;
        for (int j = 0; j < n2; j++) nrm += v[j] * v[j];
        nrm = std::sqrt(nrm);
        if (nrm < 1e-8) throw std::runtime_error("wavelet construction failed");
        for (int j = 0; j < n2; j++) b.F[(n + g) * n2 + j] = v[j] / nrm;
    }
    // Cross-correlation coefficients on [0,1].
    b.C.assign(n * n * n2, 0.0);
    {
        const Rule &ro = cached_rule(n2 + 2);
        const Rule &ri = cached_rule(n + 2);
        std::vector<double> P(n2), fa(n), fb(n), psi(n * n);
        for (size_t q = 0; q < ro.x.size(); q++) {
            double u = ro.x[q];
            std::fill(psi.begin(), psi.end(), 0.0);
            for (size_t t = 0; t < ri.x.size(); t++) {
                double xi = u + (1.0 - u) * ri.x[t];
                double wt = (1.0 - u) * ri.w[t];
                b.eval(xi, fa.data());
                b.eval(xi - u, fb.data());
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++) psi[i * n + j] += wt * fa[i] * fb[j];
            }
            legendre01(n2, u, P.data());
            for (int ij = 0; ij < n * n; ij++)
                for (int p = 0; p < n2; p++) b.C[ij * n2 + p] += ro.w[q] * psi[ij] * P[p];
        }
    }
    return b;
}

// ------------------------------------------------------------------- MRA
double MRA::box_size(int depth) const { return sfac * std::ldexp(1.0, -(n0 + depth)); }
double MRA::world_lo(int d) const { return sfac * std::ldexp((double)corner[d], -n0); }
double MRA::world_hi(int d) const { return sfac * std::ldexp((double)(corner[d] + boxes[d]), -n0); }
bool MRA::in_world(int depth, const long long *l) const {
    for (int d = 0; d < 3; d++)
        if (l[d] < tlo(depth, d) || l[d] >= thi(depth, d)) return false;
    return true;
}

}  // namespace mwg
