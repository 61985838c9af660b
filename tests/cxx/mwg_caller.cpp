// A compiled C++ caller of the C ABI (include/mwgreen.h) through the C++
// layer include/mwgreen.hpp, on a stand-in tree type in place of
// mrcpp::FunctionTree<3>: the boundary a C++ code such as MRChem links.
//
// It checks, on the GPU (device 0):
//  1. Coulomb form (CoulombPotential.cpp:126-131): apply(prec/||rho||, V, P,
//     rho) through the stand-in tree equals a direct mwg_apply (same grid,
//     coefficients to 1e-13 of the largest: the FP64 sums of a node's
//     segments are atomic, so their order, and the last bits, can differ),
//     and <rho|V> is within the projection + truncation budget of the
//     closed form sqrt(2a/pi) of a normalised Gaussian.
//  2. Exchange form (ExchangePotential.cpp:196-205): V_ij = P[rho_ij] for a
//     complex pair density (real and imaginary parts) with three precTrees,
//     each part equal to a direct mwg_apply_prec_trees.
//  3. HelmholtzVector::apply over three orbitals with distinct mu
//     (HelmholtzVector.cpp:141-159), one of them complex and conjugated: the
//     batched result equals per-part applies rescaled by -1/(2 pi) (real)
//     and sign/(2 pi) (imaginary, sign +1 when conjugated).
//  4. An error path: an apply with a negative precision is an argument
//     error reported through mwgreen::Error, and the context stays usable.
// Prints one line per check and "mwg_caller ok"; exit status 0 on success.
// With --link-only it only prints the library's version of the ABI symbols
// it calls and exits (no device needed).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "mwgreen.hpp"

// Stand-in for mrcpp::FunctionTree<3>: nodes keyed by (depth, lx, ly, lz),
// kept in insertion order (parents first), each with its compressed block.
struct StandInTree {
    std::vector<std::tuple<int, long long, long long, long long>> keys;
    std::vector<std::vector<double>> blocks;
};

void to_mwg(const StandInTree &src, mwgreen::Tree &dst) {
    std::vector<int> depth;
    std::vector<long long> l;
    std::vector<double> c;
    for (size_t i = 0; i < src.keys.size(); i++) {
        depth.push_back(std::get<0>(src.keys[i]));
        l.push_back(std::get<1>(src.keys[i]));
        l.push_back(std::get<2>(src.keys[i]));
        l.push_back(std::get<3>(src.keys[i]));
        c.insert(c.end(), src.blocks[i].begin(), src.blocks[i].end());
    }
    dst.set_nodes(depth, l, c);
}

void from_mwg(const mwgreen::Tree &src, StandInTree &dst) {
    std::vector<int> depth;
    std::vector<long long> l;
    std::vector<double> c;
    src.nodes(depth, l, c);
    const size_t b = src.block();
    dst.keys.clear();
    dst.blocks.clear();
    for (size_t i = 0; i < depth.size(); i++) {
        dst.keys.emplace_back(depth[i], l[3 * i], l[3 * i + 1], l[3 * i + 2]);
        dst.blocks.emplace_back(c.begin() + i * b, c.begin() + (i + 1) * b);
    }
}

static int failures = 0;

static void expect(bool ok, const char *what) {
    std::printf("%s: %s\n", ok ? "pass" : "FAIL", what);
    if (!ok) failures++;
}

// Largest |a - b| over two trees with identical grids (-1 if the grids differ).
static double max_diff(const StandInTree &a, const StandInTree &b) {
    if (a.keys != b.keys) return -1.0;
    double m = 0.0;
    for (size_t i = 0; i < a.blocks.size(); i++)
        for (size_t e = 0; e < a.blocks[i].size(); e++) m = std::fmax(m, std::fabs(a.blocks[i][e] - b.blocks[i][e]));
    return m;
}

static double max_abs(const StandInTree &a) {
    double m = 0.0;
    for (const auto &b : a.blocks)
        for (double v : b) m = std::fmax(m, std::fabs(v));
    return m;
}

// Same grid and coefficients equal to 1e-13 of the largest.
static bool same(const StandInTree &a, const StandInTree &b) {
    const double d = max_diff(a, b);
    return d >= 0.0 && d <= 1e-13 * std::fmax(max_abs(b), 1e-300);
}

int main(int argc, char **argv) {
    if (argc > 1 && std::strcmp(argv[1], "--link-only") == 0) {
        // resolved at load time: the program links against libmwgreen.so
        std::printf("linked: mwg_apply %p mwg_apply_batch %p mwg_apply_prec_trees %p\n", (void *)&mwg_apply,
                    (void *)&mwg_apply_batch, (void *)&mwg_apply_prec_trees);
        return 0;
    }
    using namespace mwgreen;
    const double pi = 3.14159265358979323846;
    try {
        Context ctx(0);
        const int corner[3] = {-1, -1, -1}, boxes[3] = {2, 2, 2};
        MRA mra(5, MWG_BASIS_INTERPOLATING, -4, corner, boxes, 1.25, 14);
        const double prec = 1e-4;
        Operator P = Operator::poisson(mra, prec);

        // 1. Coulomb form
        const double a = 10.0;
        Tree g(mra);
        g.project_gaussians(prec, false, {1.0}, {a}, {0.11, -0.23, 0.37});
        StandInTree rho, V;
        from_mwg(g, rho);
        const double rho_norm = std::sqrt(g.square_norm());
        apply(ctx, mra, prec / rho_norm, V, P, rho);
        Tree direct(mra);
        check(mwg_apply(ctx.get(), prec / rho_norm, direct.get(), P.get(), g.get(), -1, 0));
        StandInTree Vd;
        from_mwg(direct, Vd);
        expect(same(V, Vd), "Coulomb apply through the stand-in tree equals the direct call");
        Tree Vt(mra);
        to_mwg(V, Vt);
        const double E = g.dot(Vt), E_ex = std::sqrt(2.0 * a / pi);
        const double bound = (2.0 * rho_norm + 1.0) * prec * std::sqrt(Vt.square_norm());
        std::printf("  Coulomb energy %.10f, closed form %.10f, |dE| %.2e (budget %.2e), %zu nodes\n", E, E_ex,
                    std::fabs(E - E_ex), bound, V.keys.size());
        expect(std::fabs(E - E_ex) < bound, "Coulomb energy within the error budget of the closed form");

        // 2. Exchange form: complex pair density, three precTrees
        Tree pr(mra), pi_(mra), t1(mra), t2(mra), t3(mra);
        pr.project_gaussians(prec, true, {0.8, -0.3}, {6.0, 2.0}, {0.2, 0.1, -0.3, -0.5, 0.4, 0.2});
        pi_.project_gaussians(prec, true, {0.4}, {4.0}, {0.0, -0.3, 0.3});
        t1.project_gaussians(prec, true, {1.0}, {3.0}, {0.2, 0.1, -0.3});
        t2.project_gaussians(prec, true, {0.7}, {5.0}, {-0.5, 0.4, 0.2});
        t3.project_gaussians(prec, true, {0.5}, {2.5}, {0.0, -0.3, 0.3});
        StandInTree rr, ri, pt1, pt2, pt3, vr, vi;
        from_mwg(pr, rr);
        from_mwg(pi_, ri);
        from_mwg(t1, pt1);
        from_mwg(t2, pt2);
        from_mwg(t3, pt3);
        Complex<StandInTree> rho_ij{&rr, &ri, false}, V_ij{&vr, &vi, false};
        const double prec_p = 10.0 * prec;
        exchange_potential(ctx, mra, prec_p, P, rho_ij, {&pt1, &pt2, &pt3}, V_ij);
        const mwg_tree ph[3] = {t1.get(), t2.get(), t3.get()};
        bool all_same = true;
        for (int part = 0; part < 2; part++) {
            Tree d(mra);
            check(mwg_apply_prec_trees(ctx.get(), prec_p, d.get(), P.get(), part ? pi_.get() : pr.get(), ph, 3, -1, 1));
            StandInTree sd;
            from_mwg(d, sd);
            all_same = all_same && same(part ? vi : vr, sd);
        }
        std::printf("  exchange: real part %zu nodes, imaginary part %zu nodes\n", vr.keys.size(), vi.keys.size());
        expect(all_same, "exchange V_ij = P[rho_ij] (both parts, precTrees) equals direct applies");

        // 3. HelmholtzVector over three orbitals, one complex and conjugated
        const std::vector<double> mu = {0.9, 1.4, 2.1};
        std::vector<Operator> H;
        for (double m : mu) H.push_back(Operator::helmholtz(mra, m, prec));
        Tree o0(mra), o1(mra), o2r(mra), o2i(mra);
        o0.project_gaussians(prec, true, {1.0}, {4.0}, {0.1, 0.1, 0.1});
        o1.project_gaussians(prec, true, {0.6, 0.4}, {8.0, 1.5}, {-0.2, 0.3, 0.0, 0.4, -0.4, 0.1});
        o2r.project_gaussians(prec, true, {0.9}, {3.0}, {0.3, -0.1, -0.2});
        o2i.project_gaussians(prec, true, {-0.5}, {6.0}, {0.1, 0.2, -0.3});
        std::vector<StandInTree> in(4), out(4);
        from_mwg(o0, in[0]);
        from_mwg(o1, in[1]);
        from_mwg(o2r, in[2]);
        from_mwg(o2i, in[3]);
        std::vector<Complex<StandInTree>> phi = {{&in[0], nullptr, false}, {&in[1], nullptr, false},
                                                 {&in[2], &in[3], true}};
        std::vector<Complex<StandInTree>> res = {{&out[0], nullptr, false}, {&out[1], nullptr, false},
                                                 {&out[2], &out[3], false}};
        helmholtz_vector(ctx, mra, prec, {&H[0], &H[1], &H[2]}, phi, res);
        const Tree *src[4] = {&o0, &o1, &o2r, &o2i};
        const int opi[4] = {0, 1, 2, 2};
        const double sc[4] = {-1.0 / (2.0 * pi), -1.0 / (2.0 * pi), -1.0 / (2.0 * pi), 1.0 / (2.0 * pi)};
        double worst = 0.0;
        for (int p = 0; p < 4; p++) {
            Tree d(mra);
            check(mwg_apply(ctx.get(), prec, d.get(), H[opi[p]].get(), src[p]->get(), -1, 1));
            d.rescale(sc[p]);
            StandInTree sd;
            from_mwg(d, sd);
            const double df = max_diff(out[p], sd);
            worst = df < 0.0 ? 1.0 : std::fmax(worst, df / std::fmax(max_abs(sd), 1e-300));
        }
        std::printf("  HelmholtzVector batch of 4 parts: largest difference to single applies %.2e\n", worst);
        expect(worst <= 1e-13, "HelmholtzVector batch equals per-part applies with the -1/(2 pi), sign/(2 pi) scales");

        // 4. error path
        bool thrown = false;
        try {
            StandInTree bad;
            apply(ctx, mra, -1.0, bad, P, rho);
        } catch (const Error &e) {
            thrown = e.code == MWG_ERR_ARG;
            std::printf("  negative precision: %s\n", e.what());
        }
        expect(thrown, "an invalid apply throws mwgreen::Error with MWG_ERR_ARG");
        StandInTree again;
        apply(ctx, mra, prec / rho_norm, again, P, rho);
        expect(same(again, V), "the context still applies after the error");
    } catch (const std::exception &e) {
        std::printf("FAIL: exception %s\n", e.what());
        return 2;
    }
    if (failures) {
        std::printf("mwg_caller: %d check(s) failed\n", failures);
        return 1;
    }
    std::printf("mwg_caller ok\n");
    return 0;
}
