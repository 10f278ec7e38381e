// ildl_factor.cpp -- host multilevel inverse-based incomplete LDL^T (see
// ildl_factor.hpp for the reference routines each step restates).
#include "ildl_factor.hpp"
#include "matching.hpp"
#include "ordering.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace jdhip {

SymCsr sym_from_upper(int n, const int *ia, const int *ja, const double *a, double shift) {
    SymCsr S;
    S.n = n;
    std::vector<int> cnt(n + 1, 1);  // reserve the diagonal slot
    cnt[n] = 0;
    std::vector<int> rc(n, 1);
    for (int i = 0; i < n; ++i)
        for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
            int j = ja[k] - 1;
            if (j != i) {
                rc[i]++;
                rc[j]++;
            }
        }
    S.ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) S.ptr[i + 1] = S.ptr[i] + rc[i];
    S.idx.assign(S.ptr[n], 0);
    S.val.assign(S.ptr[n], 0.0);
    std::vector<int> fill(S.ptr.begin(), S.ptr.end() - 1);
    for (int i = 0; i < n; ++i) {
        S.idx[fill[i]] = i;
        S.val[fill[i]] = -shift;
        fill[i]++;
    }
    for (int i = 0; i < n; ++i)
        for (int k = ia[i] - 1; k < ia[i + 1] - 1; ++k) {
            int j = ja[k] - 1;
            const double v = a ? a[k] : 0.0;  // a == nullptr: pattern only
            if (j == i) {
                S.val[S.ptr[i]] += v;
            } else {
                S.idx[fill[i]] = j;
                S.val[fill[i]++] = v;
                S.idx[fill[j]] = i;
                S.val[fill[j]++] = v;
            }
        }
    // sort each row by column, merging duplicates
    std::vector<std::pair<int, double>> row;
    std::vector<int> nptr(n + 1, 0);
    std::vector<int> nidx;
    std::vector<double> nval;
    nidx.reserve(S.idx.size());
    nval.reserve(S.val.size());
    for (int i = 0; i < n; ++i) {
        row.clear();
        for (int k = S.ptr[i]; k < S.ptr[i + 1]; ++k) row.emplace_back(S.idx[k], S.val[k]);
        std::sort(row.begin(), row.end(),
                  [](const auto &x, const auto &y) { return x.first < y.first; });
        for (size_t t = 0; t < row.size(); ++t) {
            if (!nidx.empty() && (int)nidx.size() > nptr[i] && nidx.back() == row[t].first)
                nval.back() += row[t].second;
            else {
                nidx.push_back(row[t].first);
                nval.push_back(row[t].second);
            }
        }
        nptr[i + 1] = (int)nidx.size();
    }
    S.ptr.swap(nptr);
    S.idx.swap(nidx);
    S.val.swap(nval);
    return S;
}

static double pow2_near(double v) {
    if (!(v > 0.0) || !std::isfinite(v)) return 1.0;
    return std::exp2(std::round(std::log2(v)));
}

// Reverse Cuthill-McKee ordering of the pattern (ordering.cpp)
static std::vector<int> rcm_order(const SymCsr &A) {
    return rcm_order(A.n, A.ptr.data(), A.idx.data());
}

namespace {

// Progress lines on stderr when JDHIP_VERBOSE is set (long host
// factorisations would otherwise run silently for minutes).
bool verbose() {
    const char *v = std::getenv("JDHIP_VERBOSE");
    return v && *v && *v != '0';
}
double since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

template <typename T>
struct LevelOut {
    IldlLevel<T> L;
    SymCsr S;
    int nC = 0;
    int ierr = 0;
    double maxcond = 1.0;
};

// One level of the Crout ILDL with deferral (sympiluc.F restated with 1x1
// pivots).  A is the level's symmetric input matrix (double values).
template <typename T>
void factor_level(const SymCsr &A, const IldlParams &prm, double condest, long &nnz_stored,
                  long nnz_limit, LevelOut<T> &out) {
    const int n = A.n;
    IldlLevel<T> &L = out.L;
    L.n = n;
    const auto t_start = std::chrono::steady_clock::now();
    // ---- scaling (powers of two) and ordering ----
    std::vector<double> s(n);
    std::vector<int> order;
    std::vector<char> pair_start(n, 0);  // position k starts a matched 2x2 block
    if (prm.matching && prm.ordering == 2) {
        // the reference's preprocessing of every level: SYMSCALE, symmetric
        // maximum weighted matching with its scaling, 1x1/2x2 blocks, AMD on
        // the compressed graph (matching.cpp)
        const std::vector<double> s1 = symscale(A);
        SymCsr A1 = A;
        for (int i = 0; i < n; ++i)
            for (int t = A.ptr[i]; t < A.ptr[i + 1]; ++t)
                A1.val[t] = s1[i] * A.val[t] * s1[A.idx[t]];
        const MatchResult mr = mwm_symmetric(A1);
        for (int i = 0; i < n; ++i)
            for (int t = A1.ptr[i]; t < A1.ptr[i + 1]; ++t)
                A1.val[t] *= mr.scale[i] * mr.scale[A1.idx[t]];
        std::vector<int> wm;
        const int nsingle = symwm_pairs(A1, mr.match, wm);
        order = pair_amd_order(A1, wm, nsingle, pair_start);
        for (int i = 0; i < n; ++i) s[i] = s1[i] * mr.scale[i];
    } else {
        for (int i = 0; i < n; ++i) {
            double m = 0.0;
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) m = std::max(m, std::fabs(A.val[k]));
            s[i] = (m > 0.0) ? pow2_near(1.0 / std::sqrt(m)) : 1.0;
        }
        if (prm.ordering == 2) order = amd_order(n, A.ptr.data(), A.idx.data());
        else if (prm.ordering == 1) order = rcm_order(A);
        else {
            order.resize(n);
            std::iota(order.begin(), order.end(), 0);
        }
    }
    std::vector<int> pos(n);
    for (int k = 0; k < n; ++k) pos[order[k]] = k;
    // ---- scaled, permuted matrix in processing numbering (CSR, T) ----
    std::vector<int> ap(n + 1, 0), ai(A.idx.size());
    std::vector<T> av(A.idx.size());
    for (int k = 0; k < n; ++k) ap[k + 1] = ap[k] + (A.ptr[order[k] + 1] - A.ptr[order[k]]);
    {
        // A's pattern and values are symmetric, so entry A(o, j) of row
        // o = order[k] is also entry (c, k) of the permuted row c = pos[j]:
        // visiting the rows k in increasing order fills every permuted row in
        // increasing column order, with no sort.  The value is formed as
        // s[row] * a * s[column] of the permuted row's own old indices.
        std::vector<int> fill(ap.begin(), ap.end() - 1);
        for (int k = 0; k < n; ++k) {
            const int o = order[k];
            for (int t = A.ptr[o]; t < A.ptr[o + 1]; ++t) {
                const int j = A.idx[t];
                const int c = pos[j];
                const int f = fill[c]++;
                if (f >= ap[c + 1]) throw std::runtime_error("ildl: level matrix pattern is not symmetric");
                ai[f] = k;
                av[f] = (T)(s[j] * A.val[t] * s[o]);
            }
        }
    }
    const double t_order = since(t_start);
    // ---- Crout loop ----
    std::vector<int8_t> status(n, 0);  // 0 pending, 1 accepted, 2 deferred
    std::vector<int> bpos(n, -1), cpos(n, -1);
    std::vector<T> d(n, T(0));
    // inverse-norm estimators of sympiluc.F: est1 (the 1-norm estimate of
    // the columns of L^-1, w(n+1:2n)), est2 (the improved estimate with its
    // own right-hand side, w(2n+1:3n)), peaks (counted essential entries per
    // column, jw(ne+1:ne+n))
    std::vector<double> est1(n, 0.0), est2(n, 0.0);
    std::vector<int> peaks(n, 0);
    // column arena (accepted columns, rows sorted).  Column p occupies
    // [cbeg, cend): its entries of deferred rows above the current row are
    // compacted into [cbeg, dend), its entries of rows not yet reached are
    // [cur, cend), and the slots between hold entries of accepted rows,
    // which no later column needs (the level output reads the row lists)
    std::vector<long> cbeg(n, 0), cend(n, 0), cur(n, 0), dend(n, 0);
    std::vector<int> crow;
    std::vector<T> cval;
    crow.reserve(ai.size() * 2);
    cval.reserve(ai.size() * 2);
    // row lists of L: rows -> (column, value)
    std::vector<long> rhead(n, -1), rtail(n, -1), rnext;
    std::vector<int> rcol, rrow;  // column and row of every stored entry
    std::vector<T> rval;
    rnext.reserve(ai.size() * 2);
    rcol.reserve(ai.size() * 2);
    rrow.reserve(ai.size() * 2);
    rval.reserve(ai.size() * 2);
    // sparse accumulators (branch-free insertion: the index is always
    // written at the end of the list and the list grows only on first touch)
    struct Acc {
        std::vector<T> v;
        std::vector<int> mark, list;
        int len = 0;
        explicit Acc(int n_) : v(n_, T(0)), mark(n_, -1), list(n_ + 1) {}
        void clear() { len = 0; }
        void add(int i, T x, int k) {
            const bool seen = mark[i] == k;
            v[i] = seen ? v[i] + x : x;
            list[len] = i;
            len += seen ? 0 : 1;
            mark[i] = k;
        }
        const int *begin() const { return list.data(); }
        const int *end() const { return list.data() + len; }
    };
    Acc A1(n), A2(n);  // column k, and column k+1 of a 2x2 pivot candidate
    std::vector<T> &wv = A1.v, &wv2 = A2.v;
    std::vector<int> &wmark = A1.mark, &wmark2 = A2.mark;
    const Acc &wlist = A1, &wlist2 = A2;
    std::vector<std::pair<int, T>> kept;
    int nB = 0, nC = 0;
    double maxcond = 1.0;
    // 2x2 pivots: off-diagonal entry of the block and the partner row
    std::vector<T> doff(n, T(0));
    std::vector<int> partner(n, -1);
    // Column j of the not yet factored part: A(:,j) minus the updates of the
    // accepted pivots (1x1, or 2x2 blocks D_pq), into accumulator 1 or 2.
    // advance: move the column pointers past the rows before j (the deferred
    // ones go to the column's deferred list); the 2x2 partner column is
    // gathered without advancing.  Returns the updated diagonal entry.
    auto gather = [&](int j, bool second) -> T {
        Acc &W = second ? A2 : A1;
        T wj = T(0);
        for (int t = ap[j]; t < ap[j + 1]; ++t) {
            const int i = ai[t];
            if (i == j) wj = av[t];
            else if (status[i] != 1) W.add(i, av[t], j);
        }
        auto apply_col = [&](int p, T coef, T ljp) {
            wj -= coef * ljp;
            if (!second) {
                // pass the rows above j: deferred ones move to the compacted
                // prefix, accepted ones are dropped from the column
                while (cur[p] < cend[p] && crow[cur[p]] < j) {
                    const long t = cur[p]++;
                    if (status[crow[t]] == 2) {
                        crow[dend[p]] = crow[t];
                        cval[dend[p]] = cval[t];
                        ++dend[p];
                    }
                }
                // rows >= j are all pending here
                for (long t = cur[p]; t < cend[p]; ++t) {
                    const int r = crow[t];
                    if (r != j) W.add(r, -coef * cval[t], j);
                }
            } else {
                // column k+1 of a 2x2 candidate: the pointers stay at row k
                for (long t = cur[p]; t < cend[p]; ++t) {
                    const int r = crow[t];
                    if (r != j && status[r] != 1) W.add(r, -coef * cval[t], j);
                }
            }
            // deferred rows (all above j), latest first
            for (long t = dend[p] - 1; t >= cbeg[p]; --t) W.add(crow[t], -coef * cval[t], j);
        };
        for (long e = rhead[j]; e >= 0; e = rnext[e]) {
            const int p = rcol[e];
            const T l0 = rval[e];
            if (partner[p] < 0) {
                apply_col(p, l0 * d[p], l0);
            } else if (partner[p] == p + 1) {
                // 2x2 block (p, p+1): coefficients D [L(j,p) L(j,p+1)]^T
                T l1 = T(0);
                const long nx = rnext[e];
                if (nx >= 0 && rcol[nx] == p + 1) {
                    l1 = rval[nx];
                    e = nx;
                }
                apply_col(p, d[p] * l0 + doff[p] * l1, l0);
                apply_col(p + 1, doff[p] * l0 + d[p + 1] * l1, l1);
            } else {
                // second column of a block whose first entry is absent in row j
                const int q = p - 1;
                apply_col(q, doff[q] * l0, T(0));
                apply_col(p, d[p] * l0, l0);
            }
        }
        return wj;
    };
    auto store_col = [&](int k, const std::vector<std::pair<int, T>> &col) {
        cbeg[k] = (long)crow.size();
        for (auto &kv : col) {
            crow.push_back(kv.first);
            cval.push_back(kv.second);
            // append to row list of kv.first (keeps column order increasing)
            rnext.push_back(-1);
            rcol.push_back(k);
            rrow.push_back(kv.first);
            rval.push_back(kv.second);
            const long id = (long)rcol.size() - 1;
            if (rtail[kv.first] < 0) rhead[kv.first] = id;
            else rnext[rtail[kv.first]] = id;
            rtail[kv.first] = id;
        }
        cend[k] = (long)crow.size();
        cur[k] = dend[k] = cbeg[k];
    };
    std::vector<std::pair<int, T>> kept2;
    std::vector<int> urows;
    for (int k = 0; k < n;) {
        A1.clear();
        const T wk = gather(k, false);
        // ---- 1x1 or 2x2 pivot (sympiluc.F:1042-1200) ----
        // 2x2 blocks are tried where the matching paired rows k, k+1; the
        // 2x2 pivot is taken when it is more diagonally dominant than the
        // 1x1 pivot would be
        int mode = 0;
        T wk1 = T(0), b = T(0);
        double a2i11 = 0, a2i12 = 0, a2i22 = 0, mum = 0;
        double lmax1 = 0.0;
        for (int i : wlist) lmax1 = std::max(lmax1, std::fabs((double)wv[i]));
        if (k + 1 < n && pair_start[k] && status[k + 1] == 0) {
            b = (wmark[k + 1] == k) ? wv[k + 1] : T(0);
            if (b != T(0)) {
                A2.clear();
                wk1 = gather(k + 1, true);
                const double A11 = (double)wk, A12 = (double)b, A22 = (double)wk1;
                const double det = A11 * A22 - A12 * A12;
                const double mup = (A11 != 0.0) ? lmax1 / std::fabs(A11) : 0.0;
                if (det != 0.0 && std::fabs(A12) > 0.1 * std::min(std::fabs(A11), std::fabs(A22))) {
                    double mx = 0.0;
                    for (int i : wlist) {
                        if (i == k + 1) continue;
                        const double w1 = (double)wv[i];
                        const double w2 = (wmark2[i] == k + 1) ? (double)wv2[i] : 0.0;
                        mx = std::max(mx, std::fabs(A22 * w1 - A12 * w2) + std::fabs(-A12 * w1 + A11 * w2));
                    }
                    for (int i : wlist2) {
                        if (i == k || wmark[i] == k) continue;
                        const double w2 = (double)wv2[i];
                        mx = std::max(mx, std::fabs(A12 * w2) + std::fabs(A11 * w2));
                    }
                    mum = mx / std::fabs(det);
                    const double mup_eff = (A11 == 0.0) ? 10.0 * (mum + 1.0) : mup;
                    if (mup_eff > mum) {
                        mode = 1;
                        a2i11 = A22 / det;
                        a2i12 = -A12 / det;
                        a2i22 = A11 / det;
                    }
                }
            }
        }
        if (mode == 1) {
            // ---- 2x2 pivot (k, k+1) ----
            // rows of U(k:k+1, :) outside the block, with both entries
            urows.clear();
            for (int i : wlist)
                if (i != k + 1) urows.push_back(i);
            for (int i : wlist2)
                if (i != k && wmark[i] != k) urows.push_back(i);
            auto u1 = [&](int i) { return (wmark[i] == k) ? (double)wv[i] : 0.0; };
            auto u2 = [&](int i) { return (wmark2[i] == k + 1) ? (double)wv2[i] : 0.0; };
            const double U11 = a2i11, U12 = a2i12, U21 = a2i12, U22 = a2i22;
            bool skip = false;
            // first estimate with the four sign pairs (sympiluc.F:1513-1600)
            double x = 0, x2 = 0;
            {
                const double sk[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
                double X[4][2], mu[4] = {0, 0, 0, 0}, nu[4] = {0, 0, 0, 0};
                for (int c = 0; c < 4; ++c) {
                    const double r1 = sk[c][0] - est1[k], r2 = sk[c][1] - est1[k + 1];
                    X[c][0] = r1 * U11 + r2 * U21;
                    X[c][1] = r1 * U12 + r2 * U22;
                }
                for (int i : urows) {
                    const double a1 = u1(i), a2 = u2(i);
                    for (int c = 0; c < 4; ++c) {
                        const double ax = std::fabs(est1[i] + a1 * X[c][0] + a2 * X[c][1]);
                        nu[c] = std::max(nu[c], ax);
                        mu[c] += ax;
                    }
                }
                // ties as the reference: (+,+), (+,-), (-,+), then (-,-)
                int cb = 3;
                if (mu[0] >= mu[3] && mu[0] >= mu[1] && mu[0] >= mu[2]) cb = 0;
                else if (mu[1] >= mu[3] && mu[1] >= mu[0] && mu[1] >= mu[2]) cb = 1;
                else if (mu[2] >= mu[3] && mu[2] >= mu[0] && mu[2] >= mu[1]) cb = 2;
                x = X[cb][0];
                x2 = X[cb][1];
                const double axp = std::max(std::fabs(sk[cb][0] - est1[k]),
                                            std::fabs(sk[cb][1] - est1[k + 1]));
                if (std::max(nu[cb], axp) > condest) skip = true;
            }
            // peaks and first-estimate update (sympiluc.F:1642-1697)
            const double e1k1_save = est1[k + 1];
            const int pk1_save = peaks[k + 1];
            {
                int m1 = peaks[k], m2 = peaks[k + 1];
                if (std::fabs(est1[k]) <= (double)m1) ++m1;
                if (std::fabs(est1[k + 1]) <= (double)m2 && !skip) ++m2;
                peaks[k] = m1;
                if (!skip) peaks[k + 1] = m2;
                if (!skip) {
                    for (int i : urows) {
                        const double yp = u1(i) * x, ym = u2(i) * x2;
                        if (std::fabs(yp) * peaks[i] >= std::fabs(est1[i]) * m1) peaks[i] = m1;
                        if (std::fabs(ym) * peaks[i] >= std::fabs(est1[i]) * m2) peaks[i] = m2;
                        est1[i] += yp + ym;
                    }
                    est1[k] = std::max(std::fabs(1.0 - est1[k]), std::fabs(-1.0 - est1[k]));
                    est1[k + 1] =
                        std::max(std::fabs(1.0 - est1[k + 1]), std::fabs(-1.0 - est1[k + 1]));
                } else {
                    est1[k] = 1.0;
                }
            }
            // improved estimate (sympiluc.F:1883-2015)
            if (!skip) {
                const double sk[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
                double X[4][2], nu[4] = {0, 0, 0, 0};
                int cnt[4] = {0, 0, 0, 0};
                for (int c = 0; c < 4; ++c) {
                    const double r1 = sk[c][0] - est2[k], r2 = sk[c][1] - est2[k + 1];
                    X[c][0] = r1 * U11 + r2 * U21;
                    X[c][1] = r1 * U12 + r2 * U22;
                }
                for (int i : urows) {
                    const double a1 = u1(i), a2 = u2(i), ax = std::fabs(est2[i]);
                    for (int c = 0; c < 4; ++c) {
                        const double ay = std::fabs(est2[i] + a1 * X[c][0] + a2 * X[c][1]);
                        nu[c] = std::max(nu[c], ay);
                        if (ay > 2.0 * ax && ay > 0.5) ++cnt[c];
                        if (2.0 * ay < ax && ax > 0.5) --cnt[c];
                    }
                }
                // order of the reference: (+,+) j, (+,-) jj, (-,+) r, else (-,-) m
                int cb = 3;
                if (cnt[0] >= cnt[3] && cnt[0] >= cnt[1] && cnt[0] >= cnt[2]) cb = 0;
                else if (cnt[1] >= cnt[3] && cnt[1] >= cnt[0] && cnt[1] >= cnt[2]) cb = 1;
                else if (cnt[2] >= cnt[3] && cnt[2] >= cnt[0] && cnt[2] >= cnt[1]) cb = 2;
                const double axp = std::max(std::fabs(sk[cb][0] - est2[k]),
                                            std::fabs(sk[cb][1] - est2[k + 1]));
                if (std::max(nu[cb], axp) > condest) skip = true;
                if (!skip) {
                    for (int i : urows) est2[i] += u1(i) * X[cb][0] + u2(i) * X[cb][1];
                    est1[k] = std::max({est1[k], std::fabs(1.0 - est2[k]), std::fabs(-1.0 - est2[k])});
                    est1[k + 1] = std::max(
                        {est1[k + 1], std::fabs(1.0 - est2[k + 1]), std::fabs(-1.0 - est2[k + 1])});
                } else {
                    est1[k] = 1.0;
                    for (int i : urows) est1[i] -= u1(i) * x + u2(i) * x2;  // undo
                    est1[k + 1] = e1k1_save;
                    peaks[k + 1] = pk1_save;
                }
            }
            if (peaks[k] > 2) est1[k] = 2.0 * est1[k] / peaks[k];
            if (est1[k] < 1.0) est1[k] = 1.0;
            if (skip) {
                // only row k is deferred; k+1 is considered on its own next
                status[k] = 2;
                cpos[k] = nC++;
                ++k;
                continue;
            }
            if (peaks[k + 1] > 2) est1[k + 1] = 2.0 * est1[k + 1] / peaks[k + 1];
            if (est1[k + 1] < 1.0) est1[k + 1] = 1.0;
            status[k] = status[k + 1] = 1;
            bpos[k] = nB++;
            bpos[k + 1] = nB++;
            d[k] = wk;
            d[k + 1] = wk1;
            doff[k] = doff[k + 1] = b;
            partner[k] = k + 1;
            partner[k + 1] = k;
            // pair dropping (sympiluc.F:2472-2538): (U(k,l), U(k+1,l)) with
            // max |U(.,l)| Umax <= droptol / maxcondest, Umax the largest
            // ||D^-1 U(k:k+1, m)||
            maxcond = std::max({maxcond, est1[k], est1[k + 1]});
            const double thr = prm.droptol / maxcond;
            kept.clear();
            kept2.clear();
            for (int i : urows) {
                const double a1 = u1(i), a2 = u2(i);
                if (std::max(std::fabs(a1), std::fabs(a2)) * mum > thr) {
                    kept.emplace_back(i, (T)(a1 * U11 + a2 * U21));
                    kept2.emplace_back(i, (T)(a1 * U12 + a2 * U22));
                }
            }
            auto by_row = [](const auto &x1, const auto &y1) { return x1.first < y1.first; };
            std::sort(kept.begin(), kept.end(), by_row);
            std::sort(kept2.begin(), kept2.end(), by_row);
            store_col(k, kept);
            store_col(k + 1, kept2);
            nnz_stored += 2 * (long)kept.size() + 2;
            if (nnz_limit > 0 && nnz_stored > 2 * nnz_limit) {
                out.ierr = -3;
                return;
            }
            k += 2;
            continue;
        }
        bool skip = !(wk != T(0)) || !std::isfinite((double)wk);
        const double dk = (double)wk;
        // --- first estimate (sympiluc.F:1465-1510): right-hand side +-1 in
        // component k, the sign giving the larger 1-norm of the updated
        // column estimates; skip k if that estimate exceeds condest
        double x1 = 0.0;
        if (!skip) {
            const double xp = (1.0 - est1[k]) / dk, xm = (-1.0 - est1[k]) / dk;
            double mup = 0, nup = 0, mum = 0, num = 0;
            for (int i : wlist) {
                const double u = (double)wv[i];
                const double a1 = std::fabs(est1[i] + xp * u), a2 = std::fabs(est1[i] + xm * u);
                nup = std::max(nup, a1);
                mup += a1;
                num = std::max(num, a2);
                mum += a2;
            }
            double axp;
            if (mup >= mum) {
                x1 = xp;
                axp = std::fabs(1.0 - est1[k]);
            } else {
                x1 = xm;
                axp = std::fabs(-1.0 - est1[k]);
                nup = num;
            }
            if (std::max(nup, axp) > condest) skip = true;
        }
        // peaks and the update of the first estimate over the whole row
        // (before dropping), sympiluc.F:1611-1640
        {
            int m = peaks[k];
            if (std::fabs(est1[k]) <= (double)m) ++m;
            peaks[k] = m;
            if (!skip) {
                for (int i : wlist) {
                    const double y = x1 * (double)wv[i];
                    if (std::fabs(y) * peaks[i] >= std::fabs(est1[i]) * m) peaks[i] = m;
                    est1[i] += y;
                }
                est1[k] = std::max(std::fabs(1.0 - est1[k]), std::fabs(-1.0 - est1[k]));
            } else {
                est1[k] = 1.0;
            }
        }
        // --- improved estimate (sympiluc.F:1700-1860): the sign that
        // seriously increases more entries than it decreases
        if (!skip) {
            const double xp = (1.0 - est2[k]) / dk, xm = (-1.0 - est2[k]) / dk;
            int jp = 0, jm = 0;
            double nup = 0, num = 0;
            for (int i : wlist) {
                const double u = (double)wv[i];
                const double ax = std::fabs(est2[i]);
                double ay = std::fabs(est2[i] + xp * u);
                nup = std::max(nup, ay);
                if (ay > 2.0 * ax && ay > 0.5) ++jp;
                if (2.0 * ay < ax && ax > 0.5) --jp;
                ay = std::fabs(est2[i] + xm * u);
                num = std::max(num, ay);
                if (ay > 2.0 * ax && ay > 0.5) ++jm;
                if (2.0 * ay < ax && ax > 0.5) --jm;
            }
            double x2, axp;
            if (jp >= jm) {
                x2 = xp;
                axp = std::fabs(1.0 - est2[k]);
            } else {
                x2 = xm;
                axp = std::fabs(-1.0 - est2[k]);
                nup = num;
            }
            if (std::max(nup, axp) > condest) skip = true;
            if (!skip) {
                for (int i : wlist) est2[i] += x2 * (double)wv[i];
                est1[k] = std::max({est1[k], std::fabs(1.0 - est2[k]), std::fabs(-1.0 - est2[k])});
            } else {
                est1[k] = 1.0;
                for (int i : wlist) est1[i] -= x1 * (double)wv[i];  // undo the first update
            }
        }
        // infinity-norm estimate from the 1-norm and the peaks (sympiluc.F:2033)
        if (peaks[k] > 2) est1[k] = 2.0 * est1[k] / peaks[k];
        if (est1[k] < 1.0) est1[k] = 1.0;
        if (skip) {
            status[k] = 2;
            cpos[k] = nC++;
            ++k;
            continue;
        }
        status[k] = 1;
        bpos[k] = nB++;
        d[k] = wk;
        // dropping (sympiluc.F:2191-2239): U(k,l) with |U(k,l)| max_m |L(m,k)|
        // <= droptol / maxcondest, maxcondest the largest estimate so far
        maxcond = std::max(maxcond, est1[k]);
        double lmax = 0.0;
        for (int i : wlist) lmax = std::max(lmax, std::fabs((double)(wv[i] / wk)));
        const double thr = prm.droptol / maxcond;
        kept.clear();
        for (int i : wlist) {
            T li = wv[i] / wk;
            if (std::fabs((double)wv[i]) * lmax > thr) kept.emplace_back(i, li);
        }
        std::sort(kept.begin(), kept.end(),
                  [](const auto &x1, const auto &y1) { return x1.first < y1.first; });
        store_col(k, kept);
        // entries produced so far (L_B and L_CB rows): a runaway guard at
        // twice the elbow space; the elbow limit itself applies to the
        // stored factor once the level is complete
        nnz_stored += (long)kept.size() + 1;
        if (nnz_limit > 0 && nnz_stored > 2 * nnz_limit) {
            out.ierr = -3;
            return;
        }
        ++k;
    }
    out.maxcond = maxcond;
    out.nC = nC;
    L.nB = nB;
    const double t_crout = since(t_start) - t_order;
    // ---- level output ----
    L.perm.assign(n, 0);
    for (int k = 0; k < n; ++k) {
        if (status[k] == 1) L.perm[bpos[k]] = order[k];
        else L.perm[nB + cpos[k]] = order[k];
    }
    L.scal.assign(n, T(1));
    for (int i = 0; i < n; ++i) L.scal[i] = (T)s[i];
    L.d.assign(nB, T(0));
    L.doff.assign(nB, T(0));
    L.dpair.assign(nB, -1);
    std::vector<int> lbcnt(nB + 1, 0), lccnt(nC + 1, 0);
    for (int k = 0; k < n; ++k) {
        if (status[k] != 1) continue;
        L.d[bpos[k]] = d[k];
        if (partner[k] >= 0) {
            L.doff[bpos[k]] = doff[k];
            L.dpair[bpos[k]] = bpos[partner[k]];
        }
    }
    // rows of L_B and L_CB: the stored entries are numbered in column
    // (acceptance) order, so one counting pass by row keeps the column
    // indices of every row ascending (the order of its row list)
    const long nent = (long)rcol.size();
    for (long e = 0; e < nent; ++e) {
        const int r = rrow[e];
        if (status[r] == 1) ++lbcnt[bpos[r] + 1];
        else ++lccnt[cpos[r] + 1];
    }
    for (int i = 0; i < nB; ++i) lbcnt[i + 1] += lbcnt[i];
    for (int i = 0; i < nC; ++i) lccnt[i + 1] += lccnt[i];
    L.lb_ptr = lbcnt;
    L.lc_ptr = lccnt;
    L.lb_idx.assign(lbcnt[nB], 0);
    L.lb_val.assign(lbcnt[nB], T(0));
    L.lc_idx.assign(lccnt[nC], 0);
    L.lc_val.assign(lccnt[nC], T(0));
    // L_CB by columns for the Schur complement
    std::vector<int> lcc_ptr(nB + 1, 0);
    {
        std::vector<int> fb(lbcnt.begin(), lbcnt.end() - 1), fc(lccnt.begin(), lccnt.end() - 1);
        for (long e = 0; e < nent; ++e) {
            const int r = rrow[e];
            const int b = bpos[rcol[e]];
            if (status[r] == 1) {
                const int f = fb[bpos[r]]++;
                L.lb_idx[f] = b;
                L.lb_val[f] = rval[e];
            } else {
                const int f = fc[cpos[r]]++;
                L.lc_idx[f] = b;
                L.lc_val[f] = rval[e];
                lcc_ptr[b + 1]++;
            }
        }
    }
    if (nC == 0) return;
    for (int b = 0; b < nB; ++b) lcc_ptr[b + 1] += lcc_ptr[b];
    std::vector<int> lcc_idx(lcc_ptr[nB]);
    std::vector<T> lcc_val(lcc_ptr[nB]);
    {
        std::vector<int> f(lcc_ptr.begin(), lcc_ptr.end() - 1);
        for (int c = 0; c < nC; ++c)
            for (int t = L.lc_ptr[c]; t < L.lc_ptr[c + 1]; ++t) {
                int b = L.lc_idx[t];
                lcc_idx[f[b]] = c;
                lcc_val[f[b]++] = L.lc_val[t];
            }
    }
    // ---- Schur complement S = A_CC - L_CB D_B L_CB^T with dropping ----
    std::vector<int> cnode(nC);
    for (int k = 0; k < n; ++k)
        if (status[k] == 2) cnode[cpos[k]] = k;
    SymCsr &S = out.S;
    S.n = nC;
    S.ptr.assign(nC + 1, 0);
    const double sdrop = prm.droptol;
    // rows of S are independent: threads take contiguous row ranges, each
    // with its own sparse accumulator, and the pieces are concatenated in
    // row order (the result does not depend on the thread count)
    struct Piece {
        std::vector<int> cnt, idx;
        std::vector<double> val;
    };
    const int nth = std::max(1, std::min(prm.threads, std::max(1, nC / 256)));
    std::vector<Piece> pieces(nth);
    auto schur_rows = [&](int tid) {
        const int c0 = (int)((long)nC * tid / nth), c1 = (int)((long)nC * (tid + 1) / nth);
        Piece &P = pieces[tid];
        P.cnt.assign(c1 - c0, 0);
        std::vector<double> sv(nC, 0.0);
        std::vector<int> smark(nC, -1), slist(nC + 1);
        std::vector<std::pair<int, double>> srow;
        for (int c = c0; c < c1; ++c) {
            int slen = 0;
            int k = cnode[c];
            auto sacc = [&](int j, double v) {
                const bool seen = smark[j] == c;
                sv[j] = seen ? sv[j] + v : v;
                slist[slen] = j;
                slen += seen ? 0 : 1;
                smark[j] = c;
            };
            for (int t = ap[k]; t < ap[k + 1]; ++t) {
                int j = ai[t];
                if (status[j] == 2) sacc(cpos[j], (double)av[t]);
            }
            sacc(c, 0.0);
            // S(c, :) -= sum_b g_b L_CB(:, b)^T with g = D L_CB(c, :)^T (D with
            // 1x1 and 2x2 blocks; the two columns of a block are adjacent)
            auto col_update = [&](int b, double g) {
                if (g == 0.0) return;
                for (int u = lcc_ptr[b]; u < lcc_ptr[b + 1]; ++u)
                    sacc(lcc_idx[u], -g * (double)lcc_val[u]);
            };
            for (int t = L.lc_ptr[c]; t < L.lc_ptr[c + 1]; ++t) {
                const int b = L.lc_idx[t];
                const double l = (double)L.lc_val[t];
                const int q = L.dpair[b];
                if (q < 0) {
                    col_update(b, l * (double)L.d[b]);
                } else if (q == b + 1) {
                    double l1 = 0.0;
                    if (t + 1 < L.lc_ptr[c + 1] && L.lc_idx[t + 1] == b + 1) {
                        l1 = (double)L.lc_val[t + 1];
                        ++t;
                    }
                    const double o = (double)L.doff[b];
                    col_update(b, (double)L.d[b] * l + o * l1);
                    col_update(b + 1, o * l + (double)L.d[b + 1] * l1);
                } else {
                    const double o = (double)L.doff[b];
                    col_update(q, o * l);
                    col_update(b, (double)L.d[b] * l);
                }
            }
            srow.clear();
            for (int u = 0; u < slen; ++u) {
                const int j = slist[u];
                if (j == c || std::fabs(sv[j]) > sdrop) srow.emplace_back(j, sv[j]);
            }
            std::sort(srow.begin(), srow.end(),
                      [](const auto &x1, const auto &y1) { return x1.first < y1.first; });
            P.cnt[c - c0] = (int)srow.size();
            for (auto &e : srow) {
                P.idx.push_back(e.first);
                P.val.push_back(e.second);
            }
        }
    };
    if (nth == 1) {
        schur_rows(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nth; ++t) pool.emplace_back(schur_rows, t);
        for (auto &th : pool) th.join();
    }
    {
        int c = 0;
        for (auto &P : pieces) {
            for (int v : P.cnt) {
                S.ptr[c + 1] = S.ptr[c] + v;
                ++c;
            }
            S.idx.insert(S.idx.end(), P.idx.begin(), P.idx.end());
            S.val.insert(S.val.end(), P.val.begin(), P.val.end());
        }
    }
    // symmetrise the pattern and values: S := (S + S^T) / 2 on the union
    {
        std::vector<int> tp(nC + 1, 0);
        for (int c = 0; c < nC; ++c)
            for (int t = S.ptr[c]; t < S.ptr[c + 1]; ++t) tp[S.idx[t] + 1]++;
        for (int c = 0; c < nC; ++c) tp[c + 1] += tp[c];
        std::vector<int> ti(tp[nC]);
        std::vector<double> tv(tp[nC]);
        std::vector<int> f(tp.begin(), tp.end() - 1);
        for (int c = 0; c < nC; ++c)
            for (int t = S.ptr[c]; t < S.ptr[c + 1]; ++t) {
                int j = S.idx[t];
                ti[f[j]] = c;
                tv[f[j]++] = S.val[t];
            }
        SymCsr U;
        U.n = nC;
        U.ptr.assign(nC + 1, 0);
        for (int c = 0; c < nC; ++c) {
            int a = S.ptr[c], ae = S.ptr[c + 1], b = tp[c], be = tp[c + 1];
            while (a < ae || b < be) {
                int ja_ = (a < ae) ? S.idx[a] : INT32_MAX;
                int jb_ = (b < be) ? ti[b] : INT32_MAX;
                double v;
                int j;
                if (ja_ == jb_) {
                    j = ja_;
                    v = 0.5 * (S.val[a++] + tv[b++]);
                } else if (ja_ < jb_) {
                    j = ja_;
                    v = 0.5 * S.val[a++];
                } else {
                    j = jb_;
                    v = 0.5 * tv[b++];
                }
                U.idx.push_back(j);
                U.val.push_back(v);
            }
            U.ptr[c + 1] = (int)U.idx.size();
        }
        S = std::move(U);
    }
    if (prm.coupling == 0) {
        // the level's coupling block as stored and applied by the reference:
        // F^T = A_CB of the scaled, permuted input (symAMGextract.c:112);
        // L_CB was needed only for the Schur complement above
        std::vector<int> fp(nC + 1, 0), fi;
        std::vector<T> fv;
        std::vector<std::pair<int, T>> row;
        for (int c = 0; c < nC; ++c) {
            const int k = cnode[c];
            row.clear();
            for (int t = ap[k]; t < ap[k + 1]; ++t) {
                const int j = ai[t];
                if (status[j] == 1 && av[t] != T(0)) row.emplace_back(bpos[j], av[t]);
            }
            std::sort(row.begin(), row.end(),
                      [](const auto &x1, const auto &y1) { return x1.first < y1.first; });
            for (auto &e : row) {
                fi.push_back(e.first);
                fv.push_back(e.second);
            }
            fp[c + 1] = (int)fi.size();
        }
        L.lc_ptr.swap(fp);
        L.lc_idx.swap(fi);
        L.lc_val.swap(fv);
    }
    if (verbose())
        std::fprintf(stderr,
                     "jdhip ildl: level n=%d nB=%d nC=%d condest=%g  ordering %.2fs  "
                     "crout %.2fs  schur %.2fs\n",
                     n, L.nB, out.nC, condest, t_order, t_crout,
                     since(t_start) - t_order - t_crout);
}

// Dense LDL^T with Bunch-Kaufman pivoting (alpha = (1+sqrt(17))/8) of the
// final level; returns the explicit inverse of P L Dc L^T P^T where Dc is D
// (indefinite) or |D| (SPD conversion, blockwise |D| for 2x2 pivots).
int dense_inverse(const SymCsr &S, bool spd, std::vector<double> &inv, int &nneg) {
    const int m = S.n;
    std::vector<double> A((size_t)m * m, 0.0);
    auto Ar = [&](int i, int j) -> double & { return A[(size_t)i * m + j]; };
    for (int i = 0; i < m; ++i)
        for (int t = S.ptr[i]; t < S.ptr[i + 1]; ++t) Ar(i, S.idx[t]) = S.val[t];
    std::vector<int> perm(m);
    std::iota(perm.begin(), perm.end(), 0);
    std::vector<int> bsize(m, 0);  // 1 or 2 at block starts
    const double alpha = (1.0 + std::sqrt(17.0)) / 8.0;
    auto sym_swap = [&](int p, int q) {
        if (p == q) return;
        for (int j = 0; j < m; ++j) std::swap(Ar(p, j), Ar(q, j));
        for (int i = 0; i < m; ++i) std::swap(Ar(i, p), Ar(i, q));
        std::swap(perm[p], perm[q]);
    };
    std::vector<double> f1(m), f2(m);  // the pivot rows' multipliers
    int k = 0;
    while (k < m) {
        double akk = std::fabs(Ar(k, k)), colmax = 0.0;
        int imax = k;
        for (int i = k + 1; i < m; ++i)
            if (std::fabs(Ar(i, k)) > colmax) {
                colmax = std::fabs(Ar(i, k));
                imax = i;
            }
        int kstep = 1, kp = k;
        if (std::max(akk, colmax) == 0.0) return k + 1;
        if (akk < alpha * colmax) {
            double rowmax = 0.0;
            for (int j = k; j < m; ++j)
                if (j != imax) rowmax = std::max(rowmax, std::fabs(Ar(imax, j)));
            if (akk >= alpha * colmax * (colmax / rowmax)) kp = k;
            else if (std::fabs(Ar(imax, imax)) >= alpha * rowmax) kp = imax;
            else {
                kp = imax;
                kstep = 2;
            }
        }
        int kk = k + kstep - 1;
        sym_swap(kk, kp);
        // The trailing updates run row by row over contiguous memory (the
        // matrix is row-major); every entry receives the same single update
        // per step as in the column-by-column order, so the result is the
        // same to the bit.
        if (kstep == 1) {
            double dk = Ar(k, k);
            for (int i = k + 1; i < m; ++i) Ar(i, k) /= dk;
            for (int j = k + 1; j < m; ++j) f1[j] = Ar(j, k) * dk;
            for (int i = k + 1; i < m; ++i) {
                const double lik = Ar(i, k);
                double *row = &A[(size_t)i * m];
                for (int j = k + 1; j < m; ++j)
                    if (f1[j] != 0.0) row[j] -= lik * f1[j];
            }
            for (int j = k + 1; j < m; ++j) Ar(k, j) = 0.0;
            bsize[k] = 1;
        } else {
            double d11 = Ar(k, k), d21 = Ar(k + 1, k), d22 = Ar(k + 1, k + 1);
            double det = d11 * d22 - d21 * d21;
            // L(:,k:k+1) = A(:,k:k+1) D^-1
            for (int i = k + 2; i < m; ++i) {
                double a1 = Ar(i, k), a2 = Ar(i, k + 1);
                double l1 = (a1 * d22 - a2 * d21) / det;
                double l2 = (a2 * d11 - a1 * d21) / det;
                Ar(i, k) = l1;
                Ar(i, k + 1) = l2;
            }
            for (int j = k + 2; j < m; ++j) {
                f1[j] = Ar(j, k) * d11 + Ar(j, k + 1) * d21;
                f2[j] = Ar(j, k) * d21 + Ar(j, k + 1) * d22;
            }
            for (int i = k + 2; i < m; ++i) {
                const double a1 = Ar(i, k), a2 = Ar(i, k + 1);
                double *row = &A[(size_t)i * m];
                for (int j = k + 2; j < m; ++j) row[j] -= a1 * f1[j] + a2 * f2[j];
            }
            for (int j = k + 2; j < m; ++j) Ar(k, j) = Ar(k + 1, j) = 0.0;
            Ar(k, k + 1) = d21;
            bsize[k] = 2;
            bsize[k + 1] = 0;
        }
        k += kstep;
    }
    // D blocks (possibly |D|) inverted
    nneg = 0;
    std::vector<double> Dinv((size_t)m * 2, 0.0);  // per row: (self, partner) entries
    for (k = 0; k < m;) {
        if (bsize[k] == 1) {
            double dk = Ar(k, k);
            if (dk < 0) ++nneg;
            if (spd) dk = std::fabs(dk);
            Dinv[2 * k] = 1.0 / dk;
            ++k;
        } else {
            double a = Ar(k, k), b = Ar(k + 1, k), c = Ar(k + 1, k + 1);
            // eigen-decomposition of [[a b][b c]]
            double tr = 0.5 * (a + c), df = 0.5 * (a - c);
            double rt = std::sqrt(df * df + b * b);
            double l1 = tr + rt, l2 = tr - rt;
            nneg += (l1 < 0) + (l2 < 0);
            if (spd) {
                double th = 0.5 * std::atan2(2 * b, a - c);
                double cs = std::cos(th), sn = std::sin(th);
                double e1 = std::fabs(l1), e2 = std::fabs(l2);
                a = cs * cs * e1 + sn * sn * e2;
                c = sn * sn * e1 + cs * cs * e2;
                b = cs * sn * (e1 - e2);
            }
            double det = a * c - b * b;
            Dinv[2 * k] = c / det;
            Dinv[2 * k + 1] = -b / det;
            Dinv[2 * (k + 1)] = a / det;
            Dinv[2 * (k + 1) + 1] = -b / det;
            k += 2;
        }
    }
    // W = L^-1 (unit lower), then inv_perm = W^T Dinv W.  L's columns are
    // copied to contiguous rows of Lt first (Lt[i][r] = L(r, i)).
    std::vector<double> Lt((size_t)m * m, 0.0);
    for (int r = 0; r < m; ++r)
        for (int i = 0; i < r; ++i) Lt[(size_t)i * m + r] = Ar(r, i);
    std::vector<double> W((size_t)m * m, 0.0);
    std::vector<double> col(m);
    for (int j = 0; j < m; ++j) {  // column j of L^-1
        std::fill(col.begin(), col.end(), 0.0);
        col[j] = 1.0;
        for (int i = j; i < m; ++i) {
            double v = col[i];
            if (v == 0.0) continue;
            int ilo = i + 1;
            if (bsize[i] == 2) ilo = i + 2;  // L(i+1,i) is zero inside a 2x2 block
            const double *li = &Lt[(size_t)i * m];
            for (int r = ilo; r < m; ++r) col[r] -= li[r] * v;
        }
        for (int i = 0; i < m; ++i) W[(size_t)i * m + j] = col[i];
    }
    std::vector<double> DW((size_t)m * m, 0.0);
    for (k = 0; k < m;) {
        if (bsize[k] == 1) {
            for (int j = 0; j < m; ++j) DW[(size_t)k * m + j] = Dinv[2 * k] * W[(size_t)k * m + j];
            ++k;
        } else {
            for (int j = 0; j < m; ++j) {
                double w1 = W[(size_t)k * m + j], w2 = W[(size_t)(k + 1) * m + j];
                DW[(size_t)k * m + j] = Dinv[2 * k] * w1 + Dinv[2 * k + 1] * w2;
                DW[(size_t)(k + 1) * m + j] = Dinv[2 * (k + 1) + 1] * w1 + Dinv[2 * (k + 1)] * w2;
            }
            k += 2;
        }
    }
    inv.assign((size_t)m * m, 0.0);
    std::vector<double> tmp((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i)
        for (int r = 0; r < m; ++r) {
            double w = W[(size_t)r * m + i];
            if (w == 0.0) continue;
            const double *dw = &DW[(size_t)r * m];
            double *t = &tmp[(size_t)i * m];
            for (int j = 0; j < m; ++j) t[j] += w * dw[j];
        }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) inv[(size_t)perm[i] * m + perm[j]] = tmp[(size_t)i * m + j];
    return 0;
}

}  // namespace

// Eigenvalues of the symmetric 2x2 block [a b; b c] (larger first).
static void eig2(double a, double b, double c, double &l1, double &l2) {
    const double tr = 0.5 * (a + c), df = 0.5 * (a - c), rt = std::sqrt(df * df + b * b);
    l1 = tr + rt;
    l2 = tr - rt;
}

// Number of negative eigenvalues of the level's block diagonal D.
template <typename T>
static int negative_pivots(const IldlLevel<T> &L) {
    int neg = 0;
    for (int b = 0; b < L.nB; ++b) {
        const int q = L.dpair.empty() ? -1 : L.dpair[b];
        if (q < 0) {
            neg += (L.d[b] < T(0));
        } else if (q == b + 1) {
            double l1, l2;
            eig2((double)L.d[b], (double)L.doff[b], (double)L.d[b + 1], l1, l2);
            neg += (l1 < 0.0) + (l2 < 0.0);
        }
    }
    return neg;
}

// D -> |D| (SPD conversion, symspd.c:37-540): |d| for a 1x1 pivot, V |L| V^T
// for a 2x2 block with eigen-decomposition V L V^T (Jacobi rotation).
template <typename T>
static void abs_pivots(IldlLevel<T> &L) {
    for (int b = 0; b < L.nB; ++b) {
        const int q = L.dpair.empty() ? -1 : L.dpair[b];
        if (q < 0) {
            if (L.d[b] < T(0)) L.d[b] = -L.d[b];
        } else if (q == b + 1) {
            const double a = (double)L.d[b], o = (double)L.doff[b], c = (double)L.d[b + 1];
            const double th = 0.5 * std::atan2(2.0 * o, a - c);
            const double cs = std::cos(th), sn = std::sin(th);
            double l1, l2;
            eig2(a, o, c, l1, l2);
            const double e1 = std::fabs(l1), e2 = std::fabs(l2);
            L.d[b] = (T)(cs * cs * e1 + sn * sn * e2);
            L.d[b + 1] = (T)(sn * sn * e1 + cs * cs * e2);
            L.doff[b] = L.doff[b + 1] = (T)(cs * sn * (e1 - e2));
        }
    }
}

template <typename T>
IldlStats ildl_factor(const SymCsr &A0, const IldlParams &prm, IldlMultilevel<T> &out,
                      long nnz_upper_ref) {
    IldlStats st;
    out = IldlMultilevel<T>();
    out.coupling = prm.coupling;
    long nnz_stored = 0;
    long limit = (prm.mem > 0.0) ? (long)(prm.mem * (double)nnz_upper_ref) : 0;
    SymCsr cur = A0;
    double condest = prm.condest;
    st.condest_used = condest;
    int nneg = 0;
    const double condest0 = prm.condest;
    for (int lev = 0; lev < prm.max_levels && cur.n > 0; ++lev) {
        // dense switch of the coarser levels (symAMGsetup.c:450): 3 nnz(S)
        // > n(n+1)/2, nnz(S) counting the stored upper triangle with the
        // diagonal; this build also caps the dense order at dense_max
        const double full = 0.5 * (double)cur.n * (cur.n + 1.0);
        const double nnz_up = 0.5 * (double)(cur.idx.size() + (size_t)cur.n);
        if (lev > 0 && cur.n <= prm.dense_max && 3.0 * nnz_up > full) break;
        LevelOut<T> lo;
        // condest doubling while more than 75 % of the pivots are deferred,
        // up to 16 x the level's starting value and 1024 x the initial one;
        // the value reached carries over to the next level
        // (symAMGsetup.c:777-809)
        const double condest1 = condest;
        long saved = nnz_stored;
        for (;;) {
            nnz_stored = saved;
            lo = LevelOut<T>();
            factor_level<T>(cur, prm, condest, nnz_stored, limit, lo);
            if (lo.ierr) {
                st.ierr = lo.ierr;
                return st;
            }
            if (!(lo.nC > 0.75 * cur.n && condest < 1024.0 * condest0 &&
                  condest < 16.0 * condest1))
                break;
            condest *= 2.0;
        }
        if (lo.nC > 0.75 * cur.n) {
            // no sensible reduction: the level is factored completely,
            // every pivot accepted (the ILDLC fallback, symAMGsetup.c:1315)
            nnz_stored = saved;
            lo = LevelOut<T>();
            factor_level<T>(cur, prm, std::numeric_limits<double>::infinity(), nnz_stored,
                            limit, lo);
            if (lo.ierr) {
                st.ierr = lo.ierr;
                return st;
            }
        }
        st.condest_used = std::max(st.condest_used, condest);
        // elbow space (MEM): the stored factor of the levels so far (L_B,
        // coupling block, pivots) must fit into MEM x nnz(A_upper)
        nnz_stored = saved + (long)lo.L.lb_idx.size() + (long)lo.L.lc_idx.size() + lo.L.nB;
        if (limit > 0 && nnz_stored > limit) {
            st.ierr = -3;
            return st;
        }
        st.level_n.push_back(cur.n);
        st.level_nB.push_back(lo.L.nB);
        nneg += negative_pivots(lo.L);
        out.levels.push_back(std::move(lo.L));
        if (lo.nC == 0) {
            cur = SymCsr();
            break;
        }
        cur = std::move(lo.S);
    }
    if (cur.n > 0) {
        if (cur.n > 8000) {
            st.ierr = -3;
            return st;
        }
        int dn = 0;
        int info = dense_inverse(cur, prm.spd, out.dinv, dn);
        if (info) {
            st.ierr = info;
            return st;
        }
        out.m = cur.n;
        nneg += dn;
        nnz_stored += (long)cur.n * (cur.n + 1) / 2;
    }
    if (prm.spd)
        for (auto &L : out.levels) abs_pivots(L);
    out.nneg = nneg;
    out.nnz_factor = nnz_stored;
    st.nlev = (int)out.levels.size() + (out.m > 0 ? 1 : 0);
    st.nnz = nnz_stored;
    st.nneg = nneg;
    return st;
}

template IldlStats ildl_factor<double>(const SymCsr &, const IldlParams &,
                                       IldlMultilevel<double> &, long);
template IldlStats ildl_factor<float>(const SymCsr &, const IldlParams &,
                                      IldlMultilevel<float> &, long);

}  // namespace jdhip
