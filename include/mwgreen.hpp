/* mwgreen.hpp: C++ layer over the C ABI of include/mwgreen.h.
 *
 * RAII handles and the call forms MRChem makes around mrcpp::apply, for a
 * caller-side tree type T.  T is adapted by two free functions that the
 * caller provides (found by argument-dependent lookup):
 *   void to_mwg(const T &src, mwgreen::Tree &dst);   // every node, parents first
 *   void from_mwg(const mwgreen::Tree &src, T &dst);  // replace dst's nodes
 * MRChem's adapter over mrcpp::FunctionTree<3> is in INTEGRATION.md; the
 * test program tests/cxx/mwg_caller.cpp adapts a stand-in tree.
 *
 * Call forms (reference call sites):
 *   apply(ctx, prec, out, H, inp, max_iter, abs_prec)
 *       mrcpp::apply, src/qmoperators/two_electron/CoulombPotential.cpp:131,
 *       src/scf_solver/HelmholtzVector.cpp:149,154
 *   apply(ctx, prec, out, P, inp, prec_trees, max_iter, abs_prec)
 *       mrcpp::apply with precTrees, ExchangePotential.cpp:200,204
 *   helmholtz_vector(ctx, prec, ops, phi, out)
 *       HelmholtzVector::apply(i, phi) for every orbital
 *       (src/scf_solver/HelmholtzVector.cpp:141-159): the real and the
 *       imaginary part of each orbital applied with absolute precision, the
 *       real part rescaled by -1/(2 pi) and the imaginary part by
 *       sign/(2 pi), sign = +1 for a conjugated orbital and -1 otherwise;
 *       all parts of all orbitals in one batch (mwg_apply_batch).
 *   exchange_potential(ctx, prec, P, rho, prec_trees, out)
 *       V_ij = P[rho_ij] for the real and the imaginary part of a pair
 *       density with the same precTrees (ExchangePotential.cpp:196-205).
 * Errors throw mwgreen::Error with the library's message (MRChem would turn
 * it into MSG_ABORT).
 */
#ifndef MWGREEN_HPP
#define MWGREEN_HPP

#include <cmath>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mwgreen.h"

namespace mwgreen {

struct Error : std::runtime_error {
    int code;
    Error(int rc, const std::string &msg) : std::runtime_error(msg), code(rc) {}
};

inline void check(int rc) {
    if (rc != MWG_OK) throw Error(rc, mwg_last_error());
}

// Owning handle of one C ABI object.
template <typename H, int (*Destroy)(H)>
class Handle {
   public:
    Handle() = default;
    explicit Handle(H h) : h_(h) {}
    Handle(const Handle &) = delete;
    Handle &operator=(const Handle &) = delete;
    Handle(Handle &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    Handle &operator=(Handle &&o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    ~Handle() {
        if (h_) (void)Destroy(h_);
    }
    H get() const { return h_; }

   protected:
    H h_ = nullptr;
};

class Context : public Handle<mwg_context, mwg_context_destroy> {
   public:
    explicit Context(int device = 0) {
        mwg_context c = nullptr;
        check(mwg_context_create(device, &c));
        h_ = c;
    }
    void set_precision(int mode, double tau = 65536.0) { check(mwg_context_set_precision(h_, mode, tau)); }
    // (host-to-device, device-to-host) bytes of tree blocks moved by calls on
    // this context; reset zeroes the counters after reading them.
    std::pair<long long, long long> transfer_bytes(bool reset = false) const {
        long long up = 0, down = 0;
        check(mwg_context_transfer_bytes(h_, &up, &down, reset ? 1 : 0));
        return {up, down};
    }
};

// MultiResolutionAnalysis<3> (src/mrenv.cpp:99-127).
class MRA : public Handle<mwg_mra, mwg_mra_destroy> {
   public:
    MRA(int order, int basis, int root_scale, const int corner[3], const int boxes[3], double scaling_factor,
        int max_depth) {
        mwg_mra m = nullptr;
        check(mwg_mra_create(order, basis, root_scale, corner, boxes, scaling_factor, max_depth, &m));
        h_ = m;
        kp1_ = order + 1;
    }
    int block() const { return 8 * kp1_ * kp1_ * kp1_; }

   private:
    int kp1_ = 0;
};

class Operator : public Handle<mwg_operator, mwg_operator_destroy> {
   public:
    explicit Operator(mwg_operator h) : Handle(h) {}
    // PoissonOperator(MRA, prec) (src/driver.cpp:1131)
    static Operator poisson(const MRA &mra, double prec) {
        mwg_operator o = nullptr;
        check(mwg_poisson_create(mra.get(), prec, &o));
        return Operator(o);
    }
    // HelmholtzOperator(MRA, mu, prec) (src/scf_solver/HelmholtzVector.cpp:144)
    static Operator helmholtz(const MRA &mra, double mu, double prec) {
        mwg_operator o = nullptr;
        check(mwg_helmholtz_create(mra.get(), mu, prec, &o));
        return Operator(o);
    }
    // All mu at once on the device of ctx (one build per SCF iteration).
    static std::vector<Operator> helmholtz_all(const Context &ctx, const MRA &mra, const std::vector<double> &mu,
                                               double prec) {
        std::vector<mwg_operator> h(mu.size(), nullptr);
        check(mwg_operators_create_device(ctx.get(), mra.get(), (int)mu.size(), mu.data(), prec, h.data()));
        std::vector<Operator> ops;
        for (mwg_operator o : h) ops.emplace_back(o);
        return ops;
    }
};

// FunctionTree<3> of the library: nodes (depth, translation) with compressed
// 8 (k+1)^3 blocks, parents first.
class Tree : public Handle<mwg_tree, mwg_tree_destroy> {
   public:
    explicit Tree(const MRA &mra) : block_(mra.block()) {
        mwg_tree t = nullptr;
        check(mwg_tree_create(mra.get(), &t));
        h_ = t;
    }
    int block() const { return block_; }
    int size() const {
        int n = 0;
        check(mwg_tree_size(h_, &n));
        return n;
    }
    void nodes(std::vector<int> &depth, std::vector<long long> &l, std::vector<double> &coefs) const {
        const int n = size();
        depth.resize(n);
        l.resize(3 * (size_t)n);
        coefs.resize((size_t)n * block_);
        check(mwg_tree_get(h_, depth.data(), l.data(), coefs.data()));
    }
    void set_nodes(const std::vector<int> &depth, const std::vector<long long> &l, const std::vector<double> &coefs) {
        check(mwg_tree_set(h_, (int)depth.size(), depth.data(), l.data(), coefs.data()));
    }
    double square_norm() const {
        double s = 0.0;
        check(mwg_tree_square_norm(h_, &s));
        return s;
    }
    double dot(const Tree &o) const {
        double s = 0.0;
        check(mwg_tree_dot(h_, o.h_, &s));
        return s;
    }
    void rescale(double f) { check(mwg_tree_rescale(h_, f)); }
    // Keep the blocks in HBM of ctx's GPU across calls (false: back to host
    // memory); the context must outlive the residency.
    void set_resident(const Context &ctx, bool on = true) { check(mwg_tree_set_resident(h_, ctx.get(), on ? 1 : 0)); }
    void project_gaussians(double prec, bool abs_prec, const std::vector<double> &coef,
                           const std::vector<double> &exponent, const std::vector<double> &center) {
        check(mwg_tree_project_gaussians(h_, prec, abs_prec ? 1 : 0, -1, (int)coef.size(), coef.data(),
                                         exponent.data(), center.data()));
    }

   private:
    int block_ = 0;
};

// mrcpp::apply(prec, out, oper, inp, maxIter, absPrec) on caller trees.
template <typename T>
void apply(const Context &ctx, const MRA &mra, double prec, T &out, const Operator &op, const T &inp, int max_iter = -1,
           bool abs_prec = false) {
    Tree ti(mra), to(mra);
    to_mwg(inp, ti);
    check(mwg_apply(ctx.get(), prec, to.get(), op.get(), ti.get(), max_iter, abs_prec ? 1 : 0));
    from_mwg(to, out);
}

// mrcpp::apply(prec, out, oper, inp, precTrees, maxIter, absPrec).
template <typename T>
void apply(const Context &ctx, const MRA &mra, double prec, T &out, const Operator &op, const T &inp,
           const std::vector<const T *> &prec_trees, int max_iter = -1, bool abs_prec = false) {
    Tree ti(mra), to(mra);
    to_mwg(inp, ti);
    std::vector<Tree> pt;
    std::vector<mwg_tree> ph;
    pt.reserve(prec_trees.size());
    for (const T *t : prec_trees) {
        pt.emplace_back(mra);
        to_mwg(*t, pt.back());
        ph.push_back(pt.back().get());
    }
    check(mwg_apply_prec_trees(ctx.get(), prec, to.get(), op.get(), ti.get(), ph.data(), (int)ph.size(), max_iter,
                               abs_prec ? 1 : 0));
    from_mwg(to, out);
}

// A complex function as MRChem's Orbital holds it: a real and an imaginary
// part (either may be absent) and the conjugation flag.
template <typename T>
struct Complex {
    T *re = nullptr, *im = nullptr;
    bool conjugate = false;
};

// HelmholtzVector::apply(i, phi) for every i (HelmholtzVector.cpp:141-159),
// with ops[i] the Helmholtz operator of mu_i: out[i] = -2 G_mu_i[phi_i], the
// real part scaled by -1/(2 pi), the imaginary part by sign/(2 pi).  Every
// part of every orbital is one entry of a single mwg_apply_batch.
template <typename T>
void helmholtz_vector(const Context &ctx, const MRA &mra, double prec, const std::vector<const Operator *> &ops,
                      const std::vector<Complex<T>> &phi, std::vector<Complex<T>> &out) {
    if (ops.size() != phi.size() || out.size() != phi.size())
        throw Error(MWG_ERR_ARG, "helmholtz_vector: one operator and one output per orbital");
    std::vector<Tree> in, res;
    std::vector<double> scale;
    std::vector<T *> dest;
    std::vector<mwg_apply_desc> desc;
    in.reserve(2 * phi.size());
    res.reserve(2 * phi.size());
    const double pi = 3.14159265358979323846;
    for (size_t i = 0; i < phi.size(); i++) {
        const T *parts[2] = {phi[i].re, phi[i].im};
        T *outs[2] = {out[i].re, out[i].im};
        for (int p = 0; p < 2; p++) {
            if (!parts[p]) continue;
            if (!outs[p]) throw Error(MWG_ERR_ARG, "helmholtz_vector: output part missing");
            in.emplace_back(mra);
            to_mwg(*parts[p], in.back());
            res.emplace_back(mra);
            mwg_apply_desc d{};
            d.out = res.back().get();
            d.op = ops[i]->get();
            d.inp = in.back().get();
            d.prec = prec;
            d.max_iter = -1;
            d.abs_prec = 1;
            desc.push_back(d);
            scale.push_back(p == 0 ? -1.0 / (2.0 * pi) : (phi[i].conjugate ? 1.0 : -1.0) / (2.0 * pi));
            dest.push_back(outs[p]);
        }
        out[i].conjugate = false;
    }
    check(mwg_apply_batch(ctx.get(), (int)desc.size(), desc.data()));
    for (size_t e = 0; e < res.size(); e++) {
        res[e].rescale(scale[e]);
        from_mwg(res[e], *dest[e]);
    }
}

// V_ij = P[rho_ij] for both parts of a complex pair density with the same
// precTrees, absolute precision (ExchangePotential.cpp:196-205).
template <typename T>
void exchange_potential(const Context &ctx, const MRA &mra, double prec, const Operator &P, const Complex<T> &rho,
                        const std::vector<const T *> &prec_trees, Complex<T> &out) {
    const T *parts[2] = {rho.re, rho.im};
    T *outs[2] = {out.re, out.im};
    for (int p = 0; p < 2; p++) {
        if (!parts[p]) continue;
        if (!outs[p]) throw Error(MWG_ERR_ARG, "exchange_potential: output part missing");
        apply(ctx, mra, prec, *outs[p], P, *parts[p], prec_trees, -1, true);
    }
    out.conjugate = rho.conjugate;
}

}  // namespace mwgreen

#endif /* MWGREEN_HPP */
