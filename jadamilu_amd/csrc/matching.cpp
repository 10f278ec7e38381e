// matching.cpp -- symmetric maximum weighted matching, scaling and the
// compressed-graph ordering of one level of the multilevel factor.
//
// Restates the preprocessing the reference applies to every level
// (DSYMperm_mc64_amd, src/ilupack/sympermMC64amd.c:149-584):
//  1. SYMSCALE (symscale.F:5-135): D A D with D_j the nearest power of two of
//     sqrt(r_j / m_j), r_j = 1 / max_i |a_ij|, m_j = max_i r_i |a_ij|;
//  2. the epsilon graph |A| + |A|^T without the entries below 1e-2 times the
//     smaller of the two row maxima (CSRSetupGraph_epsilon.c, threshold
//     MC64THRESHOLD of ilupackmacros.h:1480);
//  3. a maximum-product perfect matching on it with its dual scaling
//     (the HSL MC64 job 5 call of sympermMC64amd.c:282-285).  MC64 is not
//     part of the reference; this is an own implementation of the published
//     algorithm (Duff and Koster, SIAM J. Matrix Anal. Appl. 22 (2001)):
//     costs c_ij = log max_k |a_kj| - log |a_ij|, a greedy start on the
//     zero-reduced-cost entries, then one Dijkstra shortest augmenting path
//     per unmatched row on the reduced costs, with dual updates.  The
//     symmetric scaling is the nearest power of two of sqrt(exp(u_i + v_i)),
//     v_i including the column normalisation (sympermMC64amd.c:297-318);
//  4. the cycles of the matching permutation broken into 1x1 and 2x2
//     blocks by block diagonal dominance (symwm.c:113-566);
//  5. approximate minimum degree on the graph with every 2x2 block merged
//     into one node, expanded so that the two rows of a block stay adjacent
//     (sympermMC64amd.c:380-560).
#include "matching.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <queue>
#include <vector>

#include "ordering.hpp"

namespace jdhip {

namespace {

// nearest power of two of v > 0 (rounding of log2 half away from zero, as
// the reference's conversions do)
double pow2_round(double v) {
    if (!(v > 0.0) || !std::isfinite(v)) return 1.0;
    const double f = std::log(v) / std::log(2.0);
    const int p = (f < 0.0) ? (int)(f - 0.5) : (int)(f + 0.5);
    return std::ldexp(1.0, p);
}

}  // namespace

std::vector<double> symscale(const SymCsr &A) {
    const int n = A.n;
    std::vector<double> r(n, 0.0), m(n, 0.0), d(n, 1.0);
    for (int i = 0; i < n; ++i)
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t) r[i] = std::max(r[i], std::fabs(A.val[t]));
    for (int i = 0; i < n; ++i) r[i] = (r[i] > 0.0) ? 1.0 / r[i] : 0.0;
    for (int i = 0; i < n; ++i)
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t)
            m[A.idx[t]] = std::max(m[A.idx[t]], r[i] * std::fabs(A.val[t]));
    for (int j = 0; j < n; ++j)
        d[j] = (r[j] > 0.0 && m[j] > 0.0) ? pow2_round(std::sqrt(r[j] / m[j])) : 1.0;
    return d;
}

MatchResult mwm_symmetric(const SymCsr &A) {
    const int n = A.n;
    MatchResult R;
    R.match.assign(n, -1);
    R.scale.assign(n, 1.0);
    // epsilon graph: keep the diagonal and |a_ij| >= 1e-2 min(rmax_i, rmax_j)
    std::vector<double> rmax(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t)
            rmax[i] = std::max(rmax[i], std::fabs(A.val[t]));
    std::vector<int> gp(n + 1, 0), gi;
    std::vector<double> ga;
    gi.reserve(A.idx.size());
    ga.reserve(A.idx.size());
    for (int i = 0; i < n; ++i) {
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t) {
            const int j = A.idx[t];
            const double v = std::fabs(A.val[t]);
            if (v == 0.0) continue;
            if (j != i && v < 1e-2 * std::min(rmax[i], rmax[j])) continue;
            gi.push_back(j);
            ga.push_back(v);
        }
        gp[i + 1] = (int)gi.size();
    }
    // column maxima (the graph is symmetric: row maxima of the kept entries)
    std::vector<double> cmax(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int t = gp[i]; t < gp[i + 1]; ++t) cmax[gi[t]] = std::max(cmax[gi[t]], ga[t]);
    // costs c_ij = log cmax_j - log |a_ij| >= 0, stored per row entry
    std::vector<double> cost(gi.size());
    for (int i = 0; i < n; ++i)
        for (int t = gp[i]; t < gp[i + 1]; ++t) cost[t] = std::log(cmax[gi[t]]) - std::log(ga[t]);
    const double inf = std::numeric_limits<double>::infinity();
    std::vector<double> u(n, 0.0), v(n, inf);
    // dual start: v_j = min_i c_ij, u_i = min_j (c_ij - v_j)
    for (int i = 0; i < n; ++i)
        for (int t = gp[i]; t < gp[i + 1]; ++t) v[gi[t]] = std::min(v[gi[t]], cost[t]);
    for (int j = 0; j < n; ++j)
        if (!std::isfinite(v[j])) v[j] = 0.0;  // empty column: structurally singular
    for (int i = 0; i < n; ++i) {
        double m = inf;
        for (int t = gp[i]; t < gp[i + 1]; ++t) m = std::min(m, cost[t] - v[gi[t]]);
        u[i] = std::isfinite(m) ? m : 0.0;
    }
    // greedy matching on tight entries (reduced cost zero)
    std::vector<int> rowof(n, -1);  // column -> matched row
    std::vector<int> &colof = R.match;  // row -> matched column
    for (int i = 0; i < n; ++i)
        for (int t = gp[i]; t < gp[i + 1]; ++t) {
            const int j = gi[t];
            if (rowof[j] < 0 && cost[t] - u[i] - v[j] <= 1e-14 * (1.0 + std::fabs(cost[t]))) {
                rowof[j] = i;
                colof[i] = j;
                break;
            }
        }
    // shortest augmenting paths (Dijkstra on reduced costs) for free rows
    std::vector<double> dist(n, inf);
    std::vector<int> pred(n, -1), visited_cols, done(n, 0);
    using QE = std::pair<double, int>;
    for (int r0 = 0; r0 < n; ++r0) {
        if (colof[r0] >= 0) continue;
        std::priority_queue<QE, std::vector<QE>, std::greater<QE>> pq;
        visited_cols.clear();
        auto relax = [&](int i, double di) {
            for (int t = gp[i]; t < gp[i + 1]; ++t) {
                const int j = gi[t];
                if (done[j]) continue;
                const double nd = di + cost[t] - u[i] - v[j];
                if (nd < dist[j]) {
                    if (!std::isfinite(dist[j])) visited_cols.push_back(j);
                    dist[j] = nd;
                    pred[j] = i;
                    pq.push({nd, j});
                }
            }
        };
        relax(r0, 0.0);
        int jfree = -1;
        double dfree = inf;
        std::vector<int> final_cols;
        while (!pq.empty()) {
            auto [dj, j] = pq.top();
            pq.pop();
            if (done[j] || dj > dist[j]) continue;
            done[j] = 1;
            final_cols.push_back(j);
            if (rowof[j] < 0) {
                jfree = j;
                dfree = dj;
                break;
            }
            relax(rowof[j], dj);
        }
        if (jfree >= 0) {
            // dual update on the settled columns and their rows
            for (int j : final_cols) {
                if (j == jfree) continue;
                const double delta = dfree - dist[j];
                v[j] -= delta;
                u[rowof[j]] += delta;
            }
            u[r0] += dfree;
            // augment along the predecessors
            int j = jfree;
            while (j >= 0) {
                const int i = pred[j];
                const int jn = colof[i];
                colof[i] = j;
                rowof[j] = i;
                j = (i == r0) ? -1 : jn;
            }
        }
        for (int j : visited_cols) {
            dist[j] = inf;
            pred[j] = -1;
            done[j] = 0;
        }
        for (int j : final_cols) done[j] = 0;
    }
    // rows left unmatched (structurally singular graph) take the free
    // columns in order
    {
        std::vector<int> freecols;
        for (int j = 0; j < n; ++j)
            if (rowof[j] < 0) freecols.push_back(j);
        size_t f = 0;
        for (int i = 0; i < n; ++i)
            if (colof[i] < 0) {
                colof[i] = freecols[f++];
                R.structurally_singular = true;
            }
    }
    // scaling: row i by exp(u_i), column j by exp(v_j) / cmax_j; the
    // symmetric scaling takes the geometric mean of row i and column i
    // (dual feasibility u_i + v_j <= c_ij means |a_ij| exp(u_i) exp(v_j) /
    // cmax_j <= 1, with equality on the matching)
    for (int i = 0; i < n; ++i) {
        const double lrow = u[i], lcol = (cmax[i] > 0.0) ? v[i] - std::log(cmax[i]) : 0.0;
        R.scale[i] = pow2_round(std::sqrt(std::exp(lrow + lcol)));
    }
    return R;
}

int symwm_pairs(const SymCsr &A, const std::vector<int> &match, std::vector<int> &order) {
    const int n = A.n;
    std::vector<int> p(match);  // cycles of the permutation i -> p[i]
    std::vector<double> diag(n, 0.0), offn(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t) {
            if (A.idx[t] == i) diag[i] = A.val[t];
            else offn[i] += std::fabs(A.val[t]);
        }
    auto entry = [&](int i, int j) {
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t)
            if (A.idx[t] == j) return A.val[t];
        return 0.0;
    };
    // block diagonal dominance of the 2x2 block (i, j) (symwm.c:197-230)
    auto weight2 = [&](int i, int j) {
        const double aij = entry(i, j), val = std::fabs(aij);
        double w = std::max(offn[i], offn[j]) - val;
        if (w < 0.0) w = 0.0;
        w *= std::max(std::fabs(diag[i]), std::fabs(diag[j])) + val;
        const double det = diag[i] * diag[j] - aij * aij;
        return w / (std::fabs(det) + 1e-30);
    };
    // break the cycle through l into 2-cycles starting with pair (s, p[s])
    // (after the optional singleton); k counts the nodes already placed
    auto split = [&](int s, int len, int k) {
        int i = s, next = p[i];
        while (k < len) {
            const int m = i;
            i = p[next];
            p[next] = m;
            next = p[i];
            k += 2;
        }
        return i;
    };
    std::vector<char> seen(n, 0);
    for (int l = 0; l < n; ++l) {
        if (seen[l]) continue;
        int len = 0;
        for (int i = l; !seen[i]; i = p[i]) {
            seen[i] = 1;
            ++len;
        }
        if (len <= 2) continue;
        if (len % 2 == 0) {
            double we = 0.0, wo = 0.0;
            bool odd = false;
            for (int k = 0, i = l; k < len; ++k, i = p[i]) {
                (odd ? wo : we) += weight2(i, p[i]);
                odd = !odd;
            }
            split(we <= wo ? l : p[l], len, 0);
        } else {
            // the most diagonally dominant node becomes the singleton
            int best = -1;
            double bw = 1e30;
            for (int k = 0, i = l; k < len; ++k, i = p[i])
                if (diag[i] != 0.0 && offn[i] / std::fabs(diag[i]) < bw) {
                    bw = offn[i] / std::fabs(diag[i]);
                    best = i;
                }
            if (best >= 0) {
                const int i0 = p[best];
                p[best] = best;
                split(i0, len, 1);
            } else {
                double we = 0.0, wo = 0.0;
                bool odd = false;
                int i = l;
                for (int k = 1; k < len; ++k, i = p[i]) {
                    (odd ? wo : we) += weight2(i, p[i]);
                    odd = !odd;
                }
                const int last = split(we <= wo ? l : p[l], len, 1);
                p[last] = last;
            }
        }
    }
    // singletons first (ascending), then the 2x2 blocks in order of their
    // first row (symwm.c:545-564)
    order.clear();
    for (int i = 0; i < n; ++i)
        if (p[i] == i) order.push_back(i);
    const int nsingle = (int)order.size();
    std::vector<char> placed(n, 0);
    for (int i = 0; i < n; ++i)
        if (p[i] != i && !placed[i]) {
            order.push_back(p[i]);
            order.push_back(i);
            placed[i] = placed[p[i]] = 1;
        }
    return nsingle;
}

std::vector<int> pair_amd_order(const SymCsr &A, const std::vector<int> &wm_order, int nsingle,
                                std::vector<char> &pair_start) {
    const int n = A.n;
    const int npairs = (n - nsingle) / 2;
    const int nc = nsingle + npairs;
    // node of every row in the compressed graph
    std::vector<int> node(n);
    for (int t = 0; t < n; ++t)
        node[wm_order[t]] = (t < nsingle) ? t : nsingle + (t - nsingle) / 2;
    std::vector<std::vector<int>> adj(nc);
    for (int i = 0; i < n; ++i)
        for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t) {
            const int a = node[i], b = node[A.idx[t]];
            if (a != b) adj[a].push_back(b);
        }
    std::vector<int> cp(nc + 1, 0), ci;
    for (int a = 0; a < nc; ++a) {
        auto &v = adj[a];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        ci.insert(ci.end(), v.begin(), v.end());
        cp[a + 1] = (int)ci.size();
    }
    const std::vector<int> co = amd_order(nc, cp.data(), ci.data());
    std::vector<int> order;
    order.reserve(n);
    pair_start.assign(n, 0);
    for (int k = 0; k < nc; ++k) {
        const int a = co[k];
        if (a < nsingle) {
            order.push_back(wm_order[a]);
        } else {
            const int t = nsingle + 2 * (a - nsingle);
            pair_start[order.size()] = 1;
            order.push_back(wm_order[t]);
            order.push_back(wm_order[t + 1]);
        }
    }
    return order;
}

}  // namespace jdhip
