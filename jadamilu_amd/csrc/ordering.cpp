// ordering.cpp -- approximate minimum degree ordering on the quotient graph,
// and reverse Cuthill-McKee (end of file).
//
// The reference orders every level of its multilevel ILDL with AMD
// (src/ilupack/permamd.c, src/ilupack/sympermMC64amd.c:503-509, calling
// SuiteSparse AMD 2.4.6, which is not part of the reference sources).  This
// is an implementation of the published algorithm: P. R. Amestoy, T. A. Davis,
// I. S. Duff, "An approximate minimum degree ordering algorithm", SIAM J.
// Matrix Anal. Appl. 17(4):886-905, 1996, i.e.
//   * the graph is kept as a quotient graph of variables and elements
//     (eliminated pivots), each variable's list holding its elements first;
//   * the pivot is a variable of least approximate external degree
//       d_i <= min(n - k, d_i_old + |Lp \ i|,
//                  |A_i \ i| + |Lp \ i| + sum_{e in E_i \ p} |Le \ Lp|),
//     the |Le \ Lp| being counted in one pass over Lp;
//   * elements adjacent to the pivot are absorbed, elements with Le inside
//     Lp are absorbed as well (aggressive absorption);
//   * variables whose only neighbour is the new element are eliminated with
//     the pivot (mass elimination); indistinguishable variables are merged
//     into supervariables, detected by hashing their lists;
//   * dense rows (degree > max(16, 10 sqrt(n))) are ordered last.
// Ties are broken last-in first-out in the degree buckets.
#include "ordering.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace jdhip {

std::vector<int> amd_order(int n, const int *ptr, const int *idx) {
    std::vector<int> order;
    if (n <= 0) return order;
    order.reserve(n);
    std::vector<int> pe(n), len(n), elen(n, 0), nv(n, 1), degree(n, 0);
    std::vector<long> w(n, 1);
    std::vector<int> head(n + 1, -1), next(n, -1), last(n, -1);
    std::vector<int> hhead(n, -1), hnext(n, -1), hkey(n, 0);
    std::vector<int> gnext(n, -1), gtail(n);
    std::vector<long> vstamp(n, 0);
    long vmark = 0;
    std::vector<int> iw;
    iw.reserve((size_t)ptr[n] + (size_t)ptr[n] / 5 + 2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        pe[i] = (int)iw.size();
        for (int k = ptr[i]; k < ptr[i + 1]; ++k)
            if (idx[k] != i) iw.push_back(idx[k]);
        len[i] = (int)iw.size() - pe[i];
        gtail[i] = i;
    }
    auto dl_insert = [&](int i, int d) {
        next[i] = head[d];
        last[i] = -1;
        if (head[d] != -1) last[head[d]] = i;
        head[d] = i;
        degree[i] = d;
    };
    auto dl_remove = [&](int i) {
        const int d = degree[i];
        if (next[i] != -1) last[next[i]] = last[i];
        if (last[i] != -1) next[last[i]] = next[i];
        else head[d] = next[i];
    };
    auto group_append = [&](int dst, int src) {
        gnext[gtail[dst]] = src;
        gtail[dst] = gtail[src];
    };

    const int dense = std::min(n - 2, std::max(16, (int)(10.0 * std::sqrt((double)n))));
    std::vector<int> dense_nodes;
    int nel = 0;
    for (int i = 0; i < n; ++i) {
        const int d = len[i];
        if (d == 0) {  // isolated: eliminated first
            elen[i] = -2;
            w[i] = 0;
            nv[i] = 0;
            order.push_back(i);
            ++nel;
        } else if (d > dense) {
            elen[i] = -1;
            nv[i] = 0;
            dense_nodes.push_back(i);
            ++nel;
        } else {
            dl_insert(i, d);
        }
    }
    // live-list compaction: variables (elen >= 0) and live elements
    const size_t gc_limit = std::max<size_t>(iw.capacity(), (size_t)2 * n + 64);
    auto compact = [&]() {
        std::vector<int> niw;
        niw.reserve(iw.capacity());
        for (int i = 0; i < n; ++i) {
            const bool var = elen[i] >= 0;
            const bool elem = elen[i] == -2 && w[i] != 0 && len[i] > 0;
            if (!var && !elem) continue;
            const int p0 = pe[i];
            pe[i] = (int)niw.size();
            niw.insert(niw.end(), iw.begin() + p0, iw.begin() + p0 + len[i]);
        }
        iw.swap(niw);
    };

    long mark = 2;
    int mindeg = 0;
    while (nel < n) {
        int p = -1;
        for (; mindeg <= n; ++mindeg)
            if (head[mindeg] != -1) {
                p = head[mindeg];
                break;
            }
        if (p < 0) break;  // cannot happen for a consistent graph
        dl_remove(p);
        const int elenp = elen[p];
        int nvp = nv[p];
        nel += nvp;

        // worst-case size of the new element: every source list
        size_t need = (size_t)(len[p] - elenp);
        for (int k = 0; k < elenp; ++k) need += (size_t)len[iw[pe[p] + k]];
        if (iw.size() + need > gc_limit) compact();

        // ---- construct the new element Lp ----
        nv[p] = -nvp;
        int degme = 0;
        const int pstart = (int)iw.size();
        const int pp = pe[p];
        for (int k = 0; k <= elenp; ++k) {
            int e, pj, ln;
            if (k < elenp) {
                e = iw[pp + k];
                pj = pe[e];
                ln = len[e];
            } else {
                e = p;
                pj = pp + elenp;
                ln = len[p] - elenp;
            }
            for (int t = 0; t < ln; ++t) {
                const int i = iw[pj + t];
                const int nvi = nv[i];
                if (nvi <= 0) continue;
                degme += nvi;
                nv[i] = -nvi;
                iw.push_back(i);
                dl_remove(i);
            }
            if (e != p) w[e] = 0;  // absorbed into p
        }
        const int pend = (int)iw.size();
        elen[p] = -2;
        pe[p] = pstart;
        len[p] = pend - pstart;
        degree[p] = degme;

        // ---- |Le \ Lp| for every element adjacent to Lp ----
        for (int t = pstart; t < pend; ++t) {
            const int i = iw[t];
            const int eln = elen[i];
            if (eln <= 0) continue;
            const int nvi = -nv[i];
            const long wnvi = mark - nvi;
            for (int k = pe[i]; k < pe[i] + eln; ++k) {
                const int e = iw[k];
                if (w[e] >= mark) w[e] -= nvi;
                else if (w[e] != 0) w[e] = degree[e] + wnvi;
            }
        }

        // ---- degree update, pruning, mass elimination, hashing ----
        for (int t = pstart; t < pend; ++t) {
            const int i = iw[t];
            const int p1 = pe[i], p2 = p1 + elen[i] - 1;
            int pn = p1;
            long hash = 0;
            long d = 0;
            for (int k = p1; k <= p2; ++k) {
                const int e = iw[k];
                if (w[e] == 0) continue;
                const long dext = w[e] - mark;
                if (dext > 0) {
                    d += dext;
                    iw[pn++] = e;
                    hash += e;
                } else {
                    w[e] = 0;  // aggressive absorption: Le inside Lp
                }
            }
            elen[i] = pn - p1 + 1;
            const int p3 = pn, p4 = p1 + len[i];
            for (int k = p2 + 1; k < p4; ++k) {
                const int j = iw[k];
                const int nvj = nv[j];
                if (nvj <= 0) continue;
                d += nvj;
                iw[pn++] = j;
                hash += j;
            }
            if (d == 0) {  // mass elimination with p
                const int nvi = -nv[i];
                degme -= nvi;
                nvp += nvi;
                nel += nvi;
                nv[i] = 0;
                elen[i] = -1;
                group_append(p, i);
            } else {
                degree[i] = (int)std::min<long>(degree[i], d);
                // p goes first; the list lost at least one entry, so slot pn is free
                iw[pn] = iw[p3];
                iw[p3] = iw[p1];
                iw[p1] = p;
                len[i] = pn - p1 + 1;
                const int h = (int)(hash % n);
                hkey[i] = h;
                hnext[i] = hhead[h];
                hhead[h] = i;
            }
        }
        degree[p] = degme;
        mark += (long)n + 1;

        // ---- supervariable detection ----
        for (int t = pstart; t < pend; ++t) {
            int i = iw[t];
            if (nv[i] >= 0) continue;
            const int h = hkey[i];
            int bucket = hhead[h];
            if (bucket == -1) continue;
            hhead[h] = -1;
            for (i = bucket; i != -1 && hnext[i] != -1; i = hnext[i]) {
                const int ln = len[i], eln = elen[i];
                ++vmark;
                for (int k = pe[i] + 1; k < pe[i] + ln; ++k) vstamp[iw[k]] = vmark;
                int jlast = i;
                for (int j = hnext[i]; j != -1;) {
                    bool same = len[j] == ln && elen[j] == eln;
                    for (int k = pe[j] + 1; same && k < pe[j] + ln; ++k)
                        if (vstamp[iw[k]] != vmark) same = false;
                    if (same) {
                        nv[i] += nv[j];  // both flagged negative
                        nv[j] = 0;
                        elen[j] = -1;
                        group_append(i, j);
                        j = hnext[j];
                        hnext[jlast] = j;
                    } else {
                        jlast = j;
                        j = hnext[j];
                    }
                }
            }
        }

        // ---- finalise degrees, compact Lp ----
        int q = pstart;
        for (int t = pstart; t < pend; ++t) {
            const int i = iw[t];
            const int nvi = -nv[i];
            if (nvi <= 0) continue;
            nv[i] = nvi;
            long dd = (long)degree[i] + degme - nvi;
            dd = std::min<long>(dd, (long)n - nel - nvi);
            dd = std::max<long>(dd, 0);
            dl_insert(i, (int)dd);
            mindeg = std::min(mindeg, (int)dd);
            iw[q++] = i;
        }
        nv[p] = nvp;
        len[p] = q - pstart;
        iw.resize(q);
        if (len[p] == 0) w[p] = 0;  // a root of the assembly tree
        for (int g = p; g != -1; g = gnext[g]) order.push_back(g);
    }
    for (int i : dense_nodes) order.push_back(i);
    if ((int)order.size() != n) {  // defensive: never return a non-permutation
        std::vector<char> seen(n, 0);
        std::vector<int> fixed;
        fixed.reserve(n);
        for (int i : order)
            if (i >= 0 && i < n && !seen[i]) {
                seen[i] = 1;
                fixed.push_back(i);
            }
        for (int i = 0; i < n; ++i)
            if (!seen[i]) fixed.push_back(i);
        order.swap(fixed);
    }
    return order;
}

// Reverse Cuthill-McKee ordering (George & Liu), one BFS per component from
// a pseudo-peripheral node; returns order[new] = old.
std::vector<int> rcm_order(int n, const int *ptr, const int *idx) {
    std::vector<int> deg(n), order;
    order.reserve(n);
    for (int i = 0; i < n; ++i) deg[i] = ptr[i + 1] - ptr[i];
    std::vector<char> seen(n, 0);
    std::vector<int> lvl(n, -1), q;
    q.reserve(n);
    auto bfs_far = [&](int s, int &far, int &depth) {
        q.clear();
        q.push_back(s);
        lvl[s] = 0;
        size_t h = 0;
        far = s;
        depth = 0;
        while (h < q.size()) {
            int u = q[h++];
            if (lvl[u] > depth || (lvl[u] == depth && deg[u] < deg[far])) {
                depth = lvl[u];
                far = u;
            }
            for (int k = ptr[u]; k < ptr[u + 1]; ++k) {
                int v = idx[k];
                if (lvl[v] < 0) {
                    lvl[v] = lvl[u] + 1;
                    q.push_back(v);
                }
            }
        }
        for (int u : q) lvl[u] = -1;
    };
    std::vector<int> nb;
    for (int start = 0; start < n; ++start) {
        if (seen[start]) continue;
        // pseudo-peripheral node: repeat BFS from the farthest node
        int s = start, far, depth, last = -1;
        for (int it = 0; it < 4; ++it) {
            bfs_far(s, far, depth);
            if (depth <= last) break;
            last = depth;
            s = far;
        }
        size_t head = order.size();
        order.push_back(s);
        seen[s] = 1;
        while (head < order.size()) {
            int u = order[head++];
            nb.clear();
            for (int k = ptr[u]; k < ptr[u + 1]; ++k) {
                int v = idx[k];
                if (!seen[v]) {
                    seen[v] = 1;
                    nb.push_back(v);
                }
            }
            std::sort(nb.begin(), nb.end(), [&](int x, int y) {
                return deg[x] < deg[y] || (deg[x] == deg[y] && x < y);
            });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}


}  // namespace jdhip
