// Apply plan: bands, generated input nodes and screening-derived bandwidths.
//
// Upstream MRCPP's OperApplicationCalculator::makeOperBand gathers, for each
// output node g at scale n, the input nodes f with |l_f - l_g|_inf <= w(n),
// generating missing ones by reconstruction from their deepest ancestor
// (MWTree::getNode with node generation; inp.deleteGenerated() afterwards in
// treebuilders/apply.cpp).  Here the same gathering happens once on the host
// into flat arrays the HIP kernels read: a coefficient pool (full blocks for
// real nodes, scaling part only for generated ones), per-slot component norms
// and a dense (2W+1)^3 band table per output node.
//
// The band half-width is derived from the screening rule itself, so it never
// changes which terms pass: along one direction, a translation l can pass
//     N_x(l) N_y N_z ||f_ft|| > thr
// only if N(l) * Nmax^2 * Fmax > thr_min, where Nmax bounds every block norm
// of the term at that depth and Fmax bounds every input component norm there.
#include "mw_internal.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <unordered_map>

namespace mwg {

namespace {
struct GenKey {
    int d;
    long long x, y, z;
    bool operator==(const GenKey &o) const { return d == o.d && x == o.x && y == o.y && z == o.z; }
};
struct GenKeyH {
    size_t operator()(const GenKey &k) const {
        uint64_t h = (uint64_t)k.d * 0x9E3779B97F4A7C15ULL;
        h ^= (uint64_t)k.x * 0xC2B2AE3D27D4EB4FULL + (h << 6) + (h >> 2);
        h ^= (uint64_t)k.y * 0x165667B19E3779F9ULL + (h << 6) + (h >> 2);
        h ^= (uint64_t)k.z * 0x27D4EB2F165667C5ULL + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};
}  // namespace

double oper_norm(const Operator &op, int m, int d, int c, long long l) {
    const OperDepth &od = op.terms[m].depth[d];
    if (od.L < 0 || l < -od.L || l > od.L) return 0.0;
    return od.norms[(size_t)c * (2 * od.L + 1) + (size_t)(l + od.L)];
}

void build_plan(ApplyPlan &plan, const Operator &op, const Tree &inp, const Tree &out,
                const std::vector<int> &nodes, const std::vector<double> &thr, int max_band_w, int seg_target,
                bool device_pool, const double *in_cnorm8) {
    const MRA &mra = *out.mra;
    const int n3 = mra.kp1_3(), blk = mra.block();
    const int M = op.size(), ndepth = op.max_depth + 1;
    if (inp.mra != out.mra || op.mra != out.mra) throw std::invalid_argument("apply: trees and operator must share one MRA");
    if (thr.size() != nodes.size()) throw std::invalid_argument("apply: one threshold per output node");
    // a fresh plan that keeps the previous pool's storage (refinement levels
    // reuse it instead of faulting in new pages)
    auto pool_storage = std::move(plan.in_coefs);
    plan = ApplyPlan();
    pool_storage.clear();
    plan.in_coefs = std::move(pool_storage);
    plan.kp1 = mra.kp1();
    plan.ndepth = ndepth;
    plan.M = M;

    int maxd = 0;
    for (int g : nodes) maxd = std::max(maxd, out.keys[g].depth);
    if (maxd >= ndepth) throw std::runtime_error("apply: output deeper than the operator tables");

    std::vector<double> thr_min(maxd + 1, 1e300), fmax(maxd + 1, 0.0);
    std::vector<char> has_depth(maxd + 1, 0);
    for (size_t q = 0; q < nodes.size(); q++) {
        int d = out.keys[nodes[q]].depth;
        thr_min[d] = std::min(thr_min[d], thr[q]);
        has_depth[d] = 1;
    }
    for (int i = 0; i < inp.size(); i++) {
        int d = inp.keys[i].depth;
        double nrm[8], tot = 0.0;
        if (in_cnorm8) {
            for (int t = 0; t < 8; t++) nrm[t] = in_cnorm8[(size_t)i * 8 + t];
        } else {
            component_norms(mra, inp.node(i), nrm);
        }
        for (int t = 0; t < 8; t++) {
            tot += nrm[t] * nrm[t];
            if (d <= maxd) fmax[d] = std::max(fmax[d], nrm[t]);
        }
        if (inp.child0[i] < 0)
            for (int dd = d + 1; dd <= maxd; dd++) fmax[dd] = std::max(fmax[dd], std::sqrt(tot));
    }

    plan.bw.assign((size_t)(maxd + 1) * M * 4, -1);
    plan.active_off.assign(maxd + 2, 0);
    std::vector<int> W(maxd + 1, 0);
    for (int d = 0; d <= maxd; d++) {
        plan.active_off[d] = (int)plan.active_terms.size();
        if (!has_depth[d]) continue;
        for (int m = 0; m < M; m++) {
            const OperDepth &od = op.terms[m].depth[d];
            if (od.L < 0) continue;
            // nmax bounds every block of the term at this depth, wmax every
            // wavelet-component block.  Off the root each candidate pairs at
            // least one wavelet component with the others, so an s<-s block
            // contributes at most N * nmax * wmax * Fmax.
            double nmax = 0.0, wmax = 0.0;
            for (int c = 0; c < 4; c++)
                for (long long l = -od.L; l <= od.L; l++) {
                    double nb = oper_norm(op, m, d, c, l);
                    nmax = std::max(nmax, nb);
                    if (c > 0) wmax = std::max(wmax, nb);
                }
            bool any = false;
            for (int c = 0; c < 4; c++) {
                double other = (d > 0 && c == 0) ? nmax * wmax : nmax * nmax;
                int w = -1;
                for (long long l = -od.L; l <= od.L; l++) {
                    double nb = oper_norm(op, m, d, c, l);
                    if (nb * other * fmax[d] * (1.0 + 1e-12) > thr_min[d]) w = std::max(w, (int)std::llabs(l));
                }
                plan.bw[((size_t)d * M + m) * 4 + c] = w;
                if (w >= 0) {
                    any = true;
                    W[d] = std::max(W[d], w);
                }
            }
            if (any) plan.active_terms.push_back(m);
        }
        if (W[d] > max_band_w)
            throw std::runtime_error("apply: band half-width " + std::to_string(W[d]) + " exceeds limit " +
                                     std::to_string(max_band_w));
    }
    plan.active_off[maxd + 1] = (int)plan.active_terms.size();

    // Output node metadata and band tables; slots are created on first use.
    std::unordered_map<GenKey, int, GenKeyH> slotmap;
    std::vector<GenKey> slot_keys;
    std::vector<int> slot_real;  // input-tree node or -1
    plan.out_node = nodes;
    plan.out_depth.resize(nodes.size());
    plan.out_thr = thr;
    plan.band_off.resize(nodes.size());
    plan.band_w.resize(nodes.size());
    long long total = 0;
    for (size_t q = 0; q < nodes.size(); q++) {
        int d = out.keys[nodes[q]].depth;
        int w = W[d];
        plan.out_depth[q] = d;
        plan.band_w[q] = w;
        plan.band_off[q] = total;
        total += (long long)(2 * w + 1) * (2 * w + 1) * (2 * w + 1);
    }
    plan.band.assign(total, -1);
    for (size_t q = 0; q < nodes.size(); q++) {
        const NodeKey &gk = out.keys[nodes[q]];
        int w = plan.band_w[q];
        long long p = plan.band_off[q];
        for (int dz = -w; dz <= w; dz++)
            for (int dy = -w; dy <= w; dy++)
                for (int dx = -w; dx <= w; dx++, p++) {
                    long long l[3] = {gk.l[0] + dx, gk.l[1] + dy, gk.l[2] + dz};
                    if (!mra.in_world(gk.depth, l)) continue;
                    GenKey k{gk.depth, l[0], l[1], l[2]};
                    auto ins = slotmap.emplace(k, (int)slot_keys.size());
                    if (ins.second) {
                        slot_keys.push_back(k);
                        slot_real.push_back(inp.find(gk.depth, l));
                    }
                    plan.band[p] = ins.first->second;
                }
    }

    // Coefficient pool: real slots keep the full block, generated slots the
    // scaling part only.
    const int ns = (int)slot_keys.size();
    plan.slot_off.resize(ns);
    long long pool = 0;
    for (int s = 0; s < ns; s++) {
        plan.slot_off[s] = pool;
        pool += (slot_real[s] >= 0) ? blk : n3;
    }
    plan.pool_values = pool;
    if (!device_pool) plan.in_coefs.resize(pool);  // every element is written below
    plan.in_norms.assign((size_t)ns * 8, 0.0);
    plan.in_depth.resize(ns);
    // Generated nodes (and the generated ancestors they descend from) are
    // collected first, then filled depth by depth in parallel: each is one
    // child of its parent, which is either a real node (full block) or a
    // generated node of the previous depth (scaling part).
    std::unordered_map<GenKey, int, GenKeyH> gen_index;
    std::vector<GenKey> gen_keys;
    std::vector<int> gen_parent;  // generated parent index, or -(real node + 1)
    auto need = [&](const GenKey &k0) -> int {
        // index of generated key k0, registering it and its generated ancestors
        std::vector<GenKey> path;
        GenKey a = k0;
        int parent = 0;
        bool found = false;
        while (true) {
            auto gi = gen_index.find(a);
            if (gi != gen_index.end() && !path.empty()) {
                parent = gi->second;
                found = true;
                break;
            }
            if (gi != gen_index.end()) return gi->second;
            long long l[3] = {a.x, a.y, a.z};
            int r = inp.find(a.d, l);
            if (r >= 0 && !path.empty()) {
                parent = -(r + 1);
                found = true;
                break;
            }
            if (a.d == 0) break;
            path.push_back(a);
            a = GenKey{a.d - 1, a.x >> 1, a.y >> 1, a.z >> 1};
        }
        if (!found) throw std::runtime_error("apply: input tree has no root covering a band node");
        for (int p = (int)path.size() - 1; p >= 0; p--) {
            const int idx = (int)gen_keys.size();
            gen_index.emplace(path[p], idx);
            gen_keys.push_back(path[p]);
            gen_parent.push_back(parent);
            parent = idx;
        }
        return parent;
    };
    std::vector<int> slot_gen(ns, -1);
    for (int s = 0; s < ns; s++)
        if (slot_real[s] < 0) slot_gen[s] = need(slot_keys[s]);
    const int ng = (int)gen_keys.size();
    for (int s = 0; s < ns; s++) plan.n_generated += (slot_real[s] < 0);
    if (device_pool) {
        // the pool is made on the device (dev_pool_fill): the sources of the
        // slots and the generated nodes in order of depth, each after its
        // parent; in_norms and the counts follow (finish_plan_counts)
        plan.device_pool = true;
        for (int s = 0; s < ns; s++) plan.in_depth[s] = slot_keys[s].d;
        // depth by depth, and within a depth the children of one parent
        // next to each other (the pool fill reconstructs them together)
        std::vector<int> order(ng), rank(ng);
        for (int i = 0; i < ng; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return gen_keys[x].d < gen_keys[y].d; });
        for (int i0 = 0; i0 < ng;) {
            int i1 = i0 + 1;
            while (i1 < ng && gen_keys[order[i1]].d == gen_keys[order[i0]].d) i1++;
            // parents of this depth are real nodes (key < 0) or generated
            // nodes of the depth above, already ranked
            auto pkey = [&](int g) {
                const int par = gen_parent[g];
                return par < 0 ? (long long)par : (long long)rank[par];
            };
            std::stable_sort(order.begin() + i0, order.begin() + i1, [&](int x, int y) { return pkey(x) < pkey(y); });
            for (int i = i0; i < i1; i++) rank[order[i]] = i;
            i0 = i1;
        }
        plan.gen_parent.resize(ng);
        plan.gen_child.resize(ng);
        plan.gen_level_off.assign(1, 0);
        for (int i = 0; i < ng; i++) {
            const int gi = order[i];
            const GenKey &ch = gen_keys[gi];
            const int par = gen_parent[gi];
            plan.gen_parent[i] = par < 0 ? par : rank[par];
            plan.gen_child[i] = (int)((ch.x & 1) | ((ch.y & 1) << 1) | ((ch.z & 1) << 2));
            if (i > 0 && ch.d != gen_keys[order[i - 1]].d) plan.gen_level_off.push_back(i);
        }
        plan.gen_level_off.push_back(ng);
        plan.slot_src.resize(ns);
        for (int s = 0; s < ns; s++) plan.slot_src[s] = slot_real[s] >= 0 ? slot_real[s] : -(rank[slot_gen[s]] + 1);
        return;
    }
    std::vector<double, UninitAlloc<double>> gen_s((size_t)ng * n3);  // written depth by depth below
    {
        std::vector<int> order(ng);
        for (int i = 0; i < ng; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return gen_keys[x].d < gen_keys[y].d; });
        size_t i0 = 0;
        while (i0 < order.size()) {
            size_t i1 = i0;
            while (i1 < order.size() && gen_keys[order[i1]].d == gen_keys[order[i0]].d) i1++;
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 8)
            for (long long ii = (long long)i0; ii < (long long)i1; ii++) {
                const int gi = order[ii];
                const GenKey &ch = gen_keys[gi];
                const int t = (int)((ch.x & 1) | ((ch.y & 1) << 1) | ((ch.z & 1) << 2));
                const int par = gen_parent[gi];
                double *o = gen_s.data() + (size_t)gi * n3;
                if (par < 0) mw_child_scaling(mra.basis, inp.node(-par - 1), false, t, o);
                else mw_child_scaling(mra.basis, gen_s.data() + (size_t)par * n3, true, t, o);
            }
            i0 = i1;
        }
    }
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 16)
    for (int s = 0; s < ns; s++) {
        plan.in_depth[s] = slot_keys[s].d;
        double *dst = plan.in_coefs.data() + plan.slot_off[s];
        if (slot_real[s] >= 0) {
            std::memcpy(dst, inp.node(slot_real[s]), sizeof(double) * blk);
            component_norms(mra, dst, &plan.in_norms[(size_t)s * 8]);
            continue;
        }
        std::memcpy(dst, gen_s.data() + (size_t)slot_gen[s] * n3, sizeof(double) * n3);
        double q = 0.0;
        for (int e = 0; e < n3; e++) q += dst[e] * dst[e];
        plan.in_norms[(size_t)s * 8] = std::sqrt(q);
    }
    finish_plan_counts(plan, op, seg_target);
}

void finish_plan_counts(ApplyPlan &plan, const Operator &op, int seg_target, bool fine_segments, int n_share) {
    const int M = op.size();
    const std::vector<int> &nodes = plan.out_node;

    // Exact term-application counts per (job, band position) with the
    // kernel's rule: a candidate (term m, input component ft) of band node f
    // for output component gt passes when |l_d| <= bw(m, d, c_d) in every
    // direction and ((N_x N_y) N_z) ||f_ft|| > thr.  Jobs are cut into
    // segments of about seg_target passing terms.
    plan.thr_floor = 0.0;
    for (double t : plan.out_thr)
        if (t > 0.0 && (plan.thr_floor == 0.0 || t < plan.thr_floor)) plan.thr_floor = t;
    const int nq = (int)nodes.size();
    // counts and bounds per (job q * 8 + gt, band position), flat: job jq's
    // band starts at cof[q] + gt * nb_q
    std::vector<long long> cof(nq + 1, 0);
    for (int q = 0; q < nq; q++) {
        const int side = 2 * plan.band_w[q] + 1;
        cof[q + 1] = cof[q] + 8LL * side * side * side;
    }
    // every entry is written by the count loop below
    std::vector<int, UninitAlloc<int>> counts(cof[nq]);
    std::vector<double, UninitAlloc<double>> bmin(cof[nq]), bmax(cof[nq]);
    // Per output depth (band half-width w): the norm the kernel would use per
    // active term, component and translation (tn), and for each component
    // combination cmb = cx | cy << 2 | cz << 4 and translation the products
    // ((N_x N_y) N_z) of the active terms, formed as the kernel forms them,
    // in ascending order (srt).  For a (band position, ft) pair the passing
    // terms are those with fl(b ||f_ft||) > thr; rounding is monotone in b, so
    // they are a suffix of the sorted products, found by bisection: the
    // counts and bounds are those of the full scan.
    std::vector<int> wd(plan.active_off.size(), -1);
    for (int q = 0; q < nq; q++) wd[plan.out_depth[q]] = plan.band_w[q];
    std::vector<std::vector<double>> tn_d(wd.size()), srt_d(wd.size());
    // norms per depth (small), then the sorted products in parallel over
    // (depth, cmb, translation): a refinement level has one output depth
    std::vector<std::pair<int, int>> items;  // (depth, cmb * nl3 + translation)
    for (int d = 0; d < (int)wd.size(); d++) {
        const int w = wd[d];
        if (w < 0 || d + 1 >= (int)plan.active_off.size()) continue;
        const int nl = 2 * w + 1, a0 = plan.active_off[d], a1 = plan.active_off[d + 1];
        std::vector<double> &tn = tn_d[d];
        tn.assign((size_t)(a1 - a0) * 4 * nl, 0.0);
        for (int ai = a0; ai < a1; ai++) {
            const int m = plan.active_terms[ai];
            const OperDepth &od = op.terms[m].depth[d];
            for (int c = 0; c < 4; c++) {
                const int bwc = plan.bw[((size_t)d * M + m) * 4 + c];
                for (int l = -w; l <= w; l++)
                    if (std::abs(l) <= bwc && std::abs(l) <= od.L)
                        tn[((size_t)(ai - a0) * 4 + c) * nl + (l + w)] = oper_norm(op, m, d, c, l);
            }
        }
        const int nl3 = nl * nl * nl;
        srt_d[d].assign((size_t)64 * nl3 * (a1 - a0), 0.0);
        for (int e = 0; e < 64 * nl3; e++) items.emplace_back(d, e);
    }
#pragma omp parallel for num_threads(team_size()) schedule(static)
    for (long long it = 0; it < (long long)items.size(); it++) {
        const int d = items[it].first, e = items[it].second;
        const int w = wd[d], nl = 2 * w + 1, nl3 = nl * nl * nl;
        const size_t na = (size_t)(plan.active_off[d + 1] - plan.active_off[d]);
        const int cmb = e / nl3, pos = e % nl3;
        const int lx = pos % nl, ly = (pos / nl) % nl, lz = pos / (nl * nl);
        const int cx = cmb & 3, cy = (cmb >> 2) & 3, cz = (cmb >> 4) & 3;
        const std::vector<double> &tn = tn_d[d];
        double *dst = srt_d[d].data() + (size_t)e * na;
        for (size_t ai = 0; ai < na; ai++)
            dst[ai] = (tn[(ai * 4 + cx) * nl + lx] * tn[(ai * 4 + cy) * nl + ly]) * tn[(ai * 4 + cz) * nl + lz];
        std::sort(dst, dst + na);
    }
    // Passing terms of output component gt of a job (its view jv) from input
    // component ft of band position p, and the smallest / largest passing
    // bound.
    struct JobView {
        int d, w, side, nl;
        size_t nl3, na;
        double th;
        const int *band;
        const double *srt;
    };
    auto view = [&](int q) {
        JobView v;
        v.d = plan.out_depth[q];
        v.w = plan.band_w[q];
        v.side = 2 * v.w + 1;
        v.nl = v.side;
        v.nl3 = (size_t)v.nl * v.nl * v.nl;
        v.na = (size_t)(plan.active_off[v.d + 1] - plan.active_off[v.d]);
        v.th = plan.out_thr[q];
        v.band = plan.band.data() + plan.band_off[q];
        v.srt = srt_d[v.d].data();
        return v;
    };
    const double *in_norms = plan.in_norms.data();
    auto pass_range = [in_norms](const JobView &v, int gt, int p, int ft, int &cnt, double &blo, double &bhi) {
        cnt = 0;
        const int sl = v.band[p];
        if (sl < 0 || (v.d > 0 && ft == 0 && gt == 0)) return;
        const double fn = in_norms[(size_t)sl * 8 + ft];
        if (!(fn > 0.0) || v.na == 0) return;
        const int w = v.w, nl = v.nl, side = v.side;
        const int ll[3] = {p % side - w, (p / side) % side - w, p / (side * side) - w};
        int cxyz[3];
        for (int dim = 0; dim < 3; dim++) cxyz[dim] = (((gt >> dim) & 1) << 1) | ((ft >> dim) & 1);
        const int cmb = cxyz[0] | (cxyz[1] << 2) | (cxyz[2] << 4);
        const size_t na = v.na;
        const double *b = v.srt + (cmb * v.nl3 + ((size_t)(ll[2] + w) * nl + (ll[1] + w)) * nl + (ll[0] + w)) * na;
        if (!(b[na - 1] * fn > v.th)) return;
        // the first product that passes
        size_t lo = 0, hi = na - 1;
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (b[mid] * fn > v.th)
                hi = mid;
            else
                lo = mid + 1;
        }
        cnt = (int)(na - lo);
        blo = b[lo] * fn;
        bhi = b[na - 1] * fn;
    };
    long long napps = 0;
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 4) reduction(+ : napps)
    for (int q = 0; q < nq; q++) {
        const JobView v = view(q);
        const int nb = v.side * v.side * v.side;
        for (int gt = 0; gt < 8; gt++) {
            int *cp = counts.data() + cof[q] + (long long)gt * nb;
            double *bl = bmin.data() + cof[q] + (long long)gt * nb, *bh = bmax.data() + cof[q] + (long long)gt * nb;
            for (int p = 0; p < nb; p++) {
                cp[p] = 0;
                bl[p] = HUGE_VAL;
                bh[p] = 0.0;
                if (v.band[p] < 0) continue;
                for (int ft = 0; ft < 8; ft++) {
                    int c = 0;
                    double l = 0.0, h = 0.0;
                    pass_range(v, gt, p, ft, c, l, h);
                    if (c == 0) continue;
                    cp[p] += c;
                    bl[p] = std::min(bl[p], l);
                    bh[p] = std::max(bh[p], h);
                }
                napps += cp[p];
            }
        }
    }
    plan.term_apps = napps;
    // A small launch (an adaptive apply's level) is cut finer, into about
    // 2048 segments of at least 32 terms, so that it still has work for every
    // compute unit's workgroups; launches with 2048 x seg_target passing terms
    // or more (this plan's count times n_share) keep seg_target.
    const int target_in = seg_target;
    const long long launch_apps = napps * std::max(1, n_share);
    seg_target = (int)std::min<long long>(seg_target, std::max<long long>(32, launch_apps / 2048));
    // In such a small plan a band position holding more passing terms than a
    // segment is cut inside, at input-component boundaries (an adaptive
    // apply's root level: a few output nodes with every term).
    const bool split_positions = seg_target < target_in;
    // Cut the (band position p, input component ft) pairs u = 8 p + ft of one
    // job (components gts[0..ng)) into segments of about `target` passing
    // terms, appended to out.  Jobs are cut in parallel, each thread over a
    // contiguous range of jobs into a buffer of its own, so that the buffers
    // in thread order hold the segments in job order.
    struct Seg {
        int job, p0, p1, apps;
        double lo, hi;
    };
    auto cut = [&](int q, const int *gts, int ng, int target, std::vector<Seg> &out) {
        const int w = plan.band_w[q], side = 2 * w + 1, nb = side * side * side;
        int u0 = 0, acc = 0;
        double lo = HUGE_VAL, hi = 0.0;
        // close the segment [u0, u1); segments without passing terms are
        // dropped (the output is zeroed before the launches)
        auto close = [&](int u1) {
            if (acc > 0) out.push_back(Seg{q * 8 + gts[0], u0, u1, acc, lo, hi});
            u0 = u1;
            acc = 0;
            lo = HUGE_VAL;
            hi = 0.0;
        };
        JobView v{};
        bool have_view = false;
        for (int p = 0; p < nb; p++) {
            int pc = 0;
            double pl = HUGE_VAL, ph = 0.0;
            for (int e = 0; e < ng; e++) {
                const long long jq = cof[q] + (long long)gts[e] * nb + p;
                pc += counts[jq];
                pl = std::min(pl, bmin[jq]);
                ph = std::max(ph, bmax[jq]);
            }
            if (split_positions && pc > target) {
                if (!have_view) {
                    v = view(q);
                    have_view = true;
                }
                for (int ft = 0; ft < 8; ft++) {
                    for (int e = 0; e < ng; e++) {
                        int c = 0;
                        double l = 0.0, h = 0.0;
                        pass_range(v, gts[e], p, ft, c, l, h);
                        if (c == 0) continue;
                        acc += c;
                        lo = std::min(lo, l);
                        hi = std::max(hi, h);
                    }
                    if (acc >= target) close(8 * p + ft + 1);
                }
            } else {
                acc += pc;
                lo = std::min(lo, pl);
                hi = std::max(hi, ph);
                if (acc >= target) close(8 * (p + 1));
            }
        }
        close(8 * nb);
    };
    const int fine_target = std::max(1, seg_target / 8);
    struct SegOut {
        std::vector<int> *job, *p0, *p1, *apps;
        std::vector<double> *lo, *hi;
    };
    // kind 0: fine single-component segments, 1: pair segments
    const SegOut outs[2] = {
        {&plan.fseg_job, &plan.fseg_p0, &plan.fseg_p1, &plan.fseg_apps, &plan.fseg_bmin, &plan.fseg_bmax},
        {&plan.pseg_job, &plan.pseg_p0, &plan.pseg_p1, &plan.pseg_apps, &plan.pseg_bmin, &plan.pseg_bmax}};
    const int nt = team_size();
    std::vector<std::vector<Seg>> buf(nt);
    for (int kind = fine_segments ? 0 : 1; kind < 2; kind++) {
        int used = 1;
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num(), T = omp_get_num_threads();
#pragma omp single
            used = T;
            std::vector<Seg> &mine = buf[t];
            mine.clear();
            const int q0 = (int)((long long)nq * t / T), q1 = (int)((long long)nq * (t + 1) / T);
            for (int q = q0; q < q1; q++)
                for (int gt = 0; gt < (kind == 1 ? 4 : 8); gt++) {
                    const int comps[2] = {gt, gt | 4};
                    cut(q, comps, kind == 1 ? 2 : 1, kind == 0 ? fine_target : seg_target, mine);
                }
        }
        size_t total = 0;
        for (int t = 0; t < used; t++) total += buf[t].size();
        const SegOut &o = outs[kind];
        for (auto *vv : {o.job, o.p0, o.p1, o.apps}) vv->resize(total);
        o.lo->resize(total);
        o.hi->resize(total);
        size_t k = 0;
        for (int t = 0; t < used; t++)
            for (const Seg &s : buf[t]) {
                (*o.job)[k] = s.job;
                (*o.p0)[k] = s.p0;
                (*o.p1)[k] = s.p1;
                (*o.apps)[k] = s.apps;
                (*o.lo)[k] = s.lo;
                (*o.hi)[k] = s.hi;
                k++;
            }
    }
}

// Pre-screen table of one apply: for each output depth d in [d0, d0 + nd),
// component combination cmb = cx | cy << 2 | cz << 4 and translation
// |l| <= w (w: the apply's widest band), the largest ((N_x N_y) N_z) over the
// depth's active terms, with the norms as the kernels stage them (zero outside
// a term's bandwidth).  Layout [d - d0][cmb][lz + w][ly + w][lx + w].
void build_prescreens(const std::vector<const ApplyPlan *> &pls, const std::vector<const OperFlat *> &fs,
                      std::vector<std::vector<double>> &mts, std::vector<std::pair<int, int>> &shapes) {
    const size_t np = pls.size();
    mts.assign(np, {});
    shapes.assign(np, {0, 0});
    // shapes and the flat task list (plan, depth offset, cmb)
    struct Task {
        int part, dd, cmb;
    };
    std::vector<Task> tasks;
    for (size_t ip = 0; ip < np; ip++) {
        const ApplyPlan &pl = *pls[ip];
        int w = 0, d0 = 0, d1 = -1;
        for (size_t q = 0; q < pl.out_depth.size(); q++) {
            w = std::max(w, pl.band_w[q]);
            if (d1 < 0 || pl.out_depth[q] < d0) d0 = pl.out_depth[q];
            d1 = std::max(d1, pl.out_depth[q]);
        }
        shapes[ip] = {d0, w};
        if (d1 < 0) {
            mts[ip].assign(1, 0.0);
            continue;
        }
        const int S = 2 * w + 1, nd = d1 - d0 + 1;
        mts[ip].assign((size_t)nd * 64 * S * S * S, 0.0);
        for (int dd = 0; dd < nd; dd++)
            for (int cmb = 0; cmb < 64; cmb++) tasks.push_back(Task{(int)ip, dd, cmb});
    }
#pragma omp parallel for num_threads(team_size()) schedule(dynamic, 8)
    for (long long ti = 0; ti < (long long)tasks.size(); ti++) {
        const Task tk = tasks[ti];
        const ApplyPlan &pl = *pls[tk.part];
        const OperFlat &f = *fs[tk.part];
        const int d0 = shapes[tk.part].first, w = shapes[tk.part].second;
        const int S = 2 * w + 1, S3 = S * S * S;
        const int d = d0 + tk.dd, cmb = tk.cmb;
        if (d + 1 >= (int)pl.active_off.size()) continue;
        const int cc[3] = {cmb & 3, (cmb >> 2) & 3, (cmb >> 4) & 3};
        double *dst = mts[tk.part].data() + ((size_t)tk.dd * 64 + cmb) * S3;
        std::vector<double> v(3 * (size_t)S);
        for (int ai = pl.active_off[d]; ai < pl.active_off[d + 1]; ai++) {
            const int m = pl.active_terms[ai];
            const int L = f.L[(size_t)m * f.ndepth + d];
            for (int dim = 0; dim < 3; dim++) {
                const int bwc = pl.bw[((size_t)d * pl.M + m) * 4 + cc[dim]];
                for (int l = -w; l <= w; l++)
                    v[dim * S + (l + w)] =
                        (std::abs(l) <= bwc && std::abs(l) <= L)
                            ? f.norms[f.noff[(size_t)m * f.ndepth + d] + (long long)cc[dim] * (2 * L + 1) + (l + L)]
                            : 0.0;
            }
            for (int lz = 0; lz < S; lz++)
                for (int ly = 0; ly < S; ly++)
                    for (int lx = 0; lx < S; lx++) {
                        const double b = (v[lx] * v[S + ly]) * v[2 * S + lz];
                        double &o = dst[((size_t)lz * S + ly) * S + lx];
                        if (b > o) o = b;
                    }
        }
    }
}

}  // namespace mwg
