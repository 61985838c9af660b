// Separated Gaussian expansions of the Poisson (1/r) and bound-state
// Helmholtz (exp(-mu r)/r) kernels.
//
// Upstream MRCPP builds these in operators/PoissonKernel.cpp and
// HelmholtzKernel.cpp (fetched by external/upstream/fetch_mrcpp.cmake:36-43,
// absent here).  This is an independent construction from the integral
// representation
//     exp(-mu r)/r = 2/sqrt(pi) int_R exp(t - r^2 e^{2t} - mu^2 e^{-2t}/4) dt
// discretised by the trapezoidal rule after the double-exponential change of
// variable t = s - exp(-s), which makes both tails decay doubly
// exponentially.  The step is shrunk until the measured error
// |K_fit(r) - K(r)| * r is below eps on a dense logarithmic grid of
// [r_min, r_max].
#include "mw_internal.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace mwg {

namespace {
struct Node {
    double alpha, beta;
};

// Largest value of r * exp(-alpha r^2) on [r_min, r_max].
double max_r_gauss(double alpha, double r_min, double r_max) {
    double rs = 1.0 / std::sqrt(2.0 * alpha);
    double r = std::min(r_max, std::max(r_min, rs));
    return r * std::exp(-alpha * r * r);
}

Node term_at(double s, double h, double mu) {
    double es = std::exp(-s);
    double t = s - es;
    double dt = 1.0 + es;
    double e2t = std::exp(2.0 * t);
    double expo = (mu > 0.0) ? t - 0.25 * mu * mu / e2t : t;
    Node n;
    n.alpha = e2t;
    n.beta = (2.0 / std::sqrt(M_PI)) * h * dt * std::exp(expo);
    return n;
}

std::vector<GaussTerm> make_terms(double h, double mu, double eps, double r_min, double r_max) {
    // Walk outwards from s = 0 until three successive terms contribute below
    // eps * 1e-3 anywhere in [r_min, r_max].
    std::vector<GaussTerm> up, down;
    double cut = eps * 1e-3;
    int quiet = 0;
    for (int j = 0; j < 100000 && quiet < 3; j++) {
        Node n = term_at(j * h, h, mu);
        double c = (n.beta > 0 && std::isfinite(n.alpha)) ? n.beta * max_r_gauss(n.alpha, r_min, r_max) : 0.0;
        if (c > cut) {
            quiet = 0;
            up.push_back({n.alpha, n.beta});
        } else {
            quiet++;
            if (c > 0) up.push_back({n.alpha, n.beta});  // keep, trimmed below
        }
    }
    quiet = 0;
    for (int j = -1; j > -100000 && quiet < 3; j--) {
        Node n = term_at(j * h, h, mu);
        double c = (n.beta > 0 && std::isfinite(n.alpha)) ? n.beta * max_r_gauss(n.alpha, r_min, r_max) : 0.0;
        if (c > cut) {
            quiet = 0;
            down.push_back({n.alpha, n.beta});
        } else {
            quiet++;
            if (c > 0) down.push_back({n.alpha, n.beta});
        }
    }
    std::vector<GaussTerm> all;
    for (auto it = down.rbegin(); it != down.rend(); ++it) all.push_back(*it);
    for (auto &t : up) all.push_back(t);
    // drop terms whose contribution is below the cut
    std::vector<GaussTerm> kept;
    for (auto &t : all)
        if (t.beta * max_r_gauss(t.alpha, r_min, r_max) > cut) kept.push_back(t);
    std::sort(kept.begin(), kept.end(), [](const GaussTerm &a, const GaussTerm &b) { return a.alpha < b.alpha; });
    return kept;
}
}  // namespace

double kernel_fit_error(const std::vector<GaussTerm> &t, double mu, double r_min, double r_max, int npts) {
    double err = 0.0;
    double l0 = std::log(r_min), l1 = std::log(r_max);
    for (int i = 0; i < npts; i++) {
        double r = std::exp(l0 + (l1 - l0) * i / (npts - 1));
        double f = 0.0;
        for (auto &g : t) f += g.beta * std::exp(-g.alpha * r * r);
        double ex = std::exp(-mu * r) / r;
        err = std::max(err, std::fabs(f - ex) * r);
    }
    return err;
}

std::vector<GaussTerm> fit_kernel(double mu, double eps, double r_min, double r_max, double *achieved_err) {
    if (!(eps > 0 && eps < 1)) throw std::invalid_argument("fit_kernel: eps must be in (0,1)");
    if (!(r_min > 0 && r_max > r_min)) throw std::invalid_argument("fit_kernel: need 0 < r_min < r_max");
    if (mu < 0) throw std::invalid_argument("fit_kernel: mu must be >= 0");
    double h = 0.6;
    std::vector<GaussTerm> terms;
    double err = 1e300;
    for (int it = 0; it < 60; it++) {
        terms = make_terms(h, mu, eps, r_min, r_max);
        err = kernel_fit_error(terms, mu, r_min, r_max, 3000);
        if (err <= 0.5 * eps) break;
        h *= 0.93;
    }
    if (err > eps) throw std::runtime_error("fit_kernel: requested accuracy not reached");
    if (achieved_err) *achieved_err = err;
    return terms;
}

}  // namespace mwg
