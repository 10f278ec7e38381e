// dense.cpp -- small dense linear algebra on the host (see dense.hpp).
#include "dense.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace jdhip {

double small_nrm2(int n, const double *x, int inc) {
    if (n < 1) return 0.0;
    if (n == 1) return std::fabs(x[0]);
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; ++i) {
        double v = x[(long)i * inc];
        if (v == 0.0) continue;
        double a = std::fabs(v);
        if (scale < a) {
            double r = scale / a;
            ssq = 1.0 + ssq * r * r;
            scale = a;
        } else {
            double r = a / scale;
            ssq += r * r;
        }
    }
    return scale * std::sqrt(ssq);
}

double small_dot(int n, const double *x, const double *y) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}

void symv_upper(int j, const double *a, int lda, const double *x, double *y) {
    for (int i = 0; i < j; ++i) y[i] = 0.0;
    for (int c = 0; c < j; ++c) {
        double t1 = x[c], t2 = 0.0;
        for (int r = 0; r < c; ++r) {
            y[r] += t1 * a[r + c * lda];
            t2 += a[r + c * lda] * x[r];
        }
        y[c] += t1 * a[c + c * lda] + t2;
    }
}

void syr2_upper(int j, double alpha, const double *x, const double *y, double *a, int lda) {
    for (int c = 0; c < j; ++c) {
        if (x[c] == 0.0 && y[c] == 0.0) continue;
        double t1 = alpha * y[c], t2 = alpha * x[c];
        for (int r = 0; r <= c; ++r) a[r + c * lda] += x[r] * t1 + y[r] * t2;
    }
}

// Cyclic Jacobi on a full symmetric j x j matrix m (column-major, ld j);
// eigenvectors accumulate in v.  Rotation formulas of Rutishauser.
static void jacobi(int j, std::vector<double> &m, std::vector<double> &v, double *w) {
    v.assign((size_t)j * j, 0.0);
    for (int i = 0; i < j; ++i) v[i + (size_t)i * j] = 1.0;
    auto M = [&](int r, int c) -> double & { return m[r + (size_t)c * j]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, dg = 0.0;
        for (int q = 0; q < j; ++q) {
            dg += M(q, q) * M(q, q);
            for (int p = 0; p < q; ++p) off += M(p, q) * M(p, q);
        }
        if (off == 0.0 || off <= 1e-36 * dg) break;
        for (int p = 0; p < j - 1; ++p)
            for (int q = p + 1; q < j; ++q) {
                double apq = M(p, q);
                if (apq == 0.0) continue;
                double app = M(p, p), aqq = M(q, q), g = 100.0 * std::fabs(apq);
                if (sweep > 4 && std::fabs(app) + g == std::fabs(app) &&
                    std::fabs(aqq) + g == std::fabs(aqq)) {
                    M(p, q) = M(q, p) = 0.0;
                    continue;
                }
                double th = (aqq - app) / (2.0 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c, tau = s / (1.0 + c);
                M(p, p) = app - t * apq;
                M(q, q) = aqq + t * apq;
                M(p, q) = M(q, p) = 0.0;
                for (int r = 0; r < j; ++r) {
                    if (r == p || r == q) continue;
                    double arp = M(r, p), arq = M(r, q);
                    double nrp = arp - s * (arq + tau * arp), nrq = arq + s * (arp - tau * arq);
                    M(r, p) = M(p, r) = nrp;
                    M(r, q) = M(q, r) = nrq;
                }
                for (int r = 0; r < j; ++r) {
                    double vrp = v[r + (size_t)p * j], vrq = v[r + (size_t)q * j];
                    v[r + (size_t)p * j] = vrp - s * (vrq + tau * vrp);
                    v[r + (size_t)q * j] = vrq + s * (vrp - tau * vrq);
                }
            }
    }
    for (int i = 0; i < j; ++i) w[i] = M(i, i);
    for (int i = 0; i < j - 1; ++i) {
        int k = i;
        for (int l = i + 1; l < j; ++l)
            if (w[l] < w[k]) k = l;
        if (k != i) {
            std::swap(w[i], w[k]);
            for (int r = 0; r < j; ++r) std::swap(v[r + (size_t)i * j], v[r + (size_t)k * j]);
        }
    }
}

int sym_eig(int j, double *a, int lda, double *w) {
    std::vector<double> m((size_t)j * j), v;
    for (int c = 0; c < j; ++c)
        for (int r = 0; r <= c; ++r) m[r + (size_t)c * j] = m[c + (size_t)r * j] = a[r + c * lda];
    jacobi(j, m, v, w);
    for (int c = 0; c < j; ++c)
        for (int r = 0; r < j; ++r) a[r + c * lda] = v[r + (size_t)c * j];
    return 0;
}

int sym_gen_eig(int j, double *a, int lda, double *b, int ldb, double *w) {
    // B = U^T U, upper Cholesky in place
    for (int c = 0; c < j; ++c) {
        double d = b[c + c * ldb];
        for (int k = 0; k < c; ++k) d -= b[k + c * ldb] * b[k + c * ldb];
        if (!(d > 0.0)) return j + c + 1;
        d = std::sqrt(d);
        b[c + c * ldb] = d;
        for (int cc = c + 1; cc < j; ++cc) {
            double s = b[c + cc * ldb];
            for (int k = 0; k < c; ++k) s -= b[k + c * ldb] * b[k + cc * ldb];
            b[c + cc * ldb] = s / d;
        }
    }
    std::vector<double> m((size_t)j * j), t((size_t)j * j), v;
    for (int c = 0; c < j; ++c)
        for (int r = 0; r <= c; ++r) m[r + (size_t)c * j] = m[c + (size_t)r * j] = a[r + c * lda];
    for (int c = 0; c < j; ++c)  // t = U^-T m
        for (int r = 0; r < j; ++r) {
            double s = m[r + (size_t)c * j];
            for (int k = 0; k < r; ++k) s -= b[k + r * ldb] * t[k + (size_t)c * j];
            t[r + (size_t)c * j] = s / b[r + r * ldb];
        }
    for (int r = 0; r < j; ++r)  // m = t U^-1
        for (int c = 0; c < j; ++c) {
            double s = t[r + (size_t)c * j];
            for (int k = 0; k < c; ++k) s -= b[k + c * ldb] * m[r + (size_t)k * j];
            m[r + (size_t)c * j] = s / b[c + c * ldb];
        }
    for (int c = 0; c < j; ++c)
        for (int r = 0; r < c; ++r) {
            double s = 0.5 * (m[r + (size_t)c * j] + m[c + (size_t)r * j]);
            m[r + (size_t)c * j] = m[c + (size_t)r * j] = s;
        }
    jacobi(j, m, v, w);
    for (int c = 0; c < j; ++c)  // y = U^-1 z
        for (int r = j - 1; r >= 0; --r) {
            double s = v[r + (size_t)c * j];
            for (int k = r + 1; k < j; ++k) s -= b[r + k * ldb] * a[k + c * lda];
            a[r + c * lda] = s / b[r + r * ldb];
        }
    return 0;
}

static void house(int n, double &alpha, double *x, double &tau) {
    tau = 0.0;
    if (n <= 1) return;
    double xn = small_nrm2(n - 1, x);
    if (xn == 0.0) return;
    double beta = -std::copysign(std::hypot(alpha, xn), alpha);
    tau = (beta - alpha) / beta;
    double sc = 1.0 / (alpha - beta);
    for (int i = 0; i < n - 1; ++i) x[i] *= sc;
    alpha = beta;
}

void house_qr(int m, int k, double *a, int lda, double *tau) {
    int kk = std::min(m, k);
    for (int i = 0; i < kk; ++i) {
        double *col = a + i + (long)i * lda;
        house(m - i, col[0], col + 1, tau[i]);
        if (i < k - 1) {
            double aii = col[0];
            col[0] = 1.0;
            for (int c = i + 1; c < k; ++c) {
                double *cc = a + i + (long)c * lda, w = 0.0;
                for (int r = 0; r < m - i; ++r) w += cc[r] * col[r];
                for (int r = 0; r < m - i; ++r) cc[r] -= tau[i] * w * col[r];
            }
            col[0] = aii;
        }
    }
}

void apply_q_right(int nr, int m, int nref, const double *a, int lda, const double *tau,
                   double *c, int ldc) {
    std::vector<double> v(m), w(nr);
    for (int i = 0; i < nref; ++i) {
        if (tau[i] == 0.0) continue;
        int len = m - i;
        v[0] = 1.0;
        for (int r = 1; r < len; ++r) v[r] = a[i + r + (long)i * lda];
        std::fill(w.begin(), w.end(), 0.0);
        for (int cc = 0; cc < len; ++cc) {
            const double *col = c + (long)(i + cc) * ldc;
            for (int r = 0; r < nr; ++r) w[r] += v[cc] * col[r];
        }
        for (int cc = 0; cc < len; ++cc) {
            double t = -tau[i] * v[cc];
            double *col = c + (long)(i + cc) * ldc;
            for (int r = 0; r < nr; ++r) col[r] += w[r] * t;
        }
    }
}

void apply_qt_left(int m, int nc, int nref, const double *a, int lda, const double *tau,
                   double *c, int ldc) {
    std::vector<double> v(m);
    for (int i = 0; i < nref; ++i) {
        if (tau[i] == 0.0) continue;
        int len = m - i;
        v[0] = 1.0;
        for (int r = 1; r < len; ++r) v[r] = a[i + r + (long)i * lda];
        for (int cc = 0; cc < nc; ++cc) {
            double *col = c + i + (long)cc * ldc, w = 0.0;
            for (int r = 0; r < len; ++r) w += col[r] * v[r];
            w *= -tau[i];
            for (int r = 0; r < len; ++r) col[r] += v[r] * w;
        }
    }
}

void form_q(int m, int ncol, int nref, const double *a, int lda, const double *tau,
            double *q) {
    // Q(:,0:ncol) = H_1 ... H_nref I(:,0:ncol): identity rows, then C := C Q
    // applied to the m x m identity is the same as transposed application;
    // build row by row: Q = I * Q.
    std::vector<double> full((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) full[i + (size_t)i * m] = 1.0;
    apply_q_right(m, m, nref, a, lda, tau, full.data(), m);
    for (int c = 0; c < ncol; ++c)
        for (int r = 0; r < m; ++r) q[r + (size_t)c * m] = full[r + (size_t)c * m];
}

void sortrx(int n, const double *data, int *index) {
    const int M = 9;
    std::vector<int> ix(n + 1);
    auto d = [&](int k) { return data[k - 1]; };
    for (int i = 1; i <= n; ++i) ix[i] = i;
    if (n > M) {
        int lstk[64], rstk[64], istk = 0, l = 1, r = n;
        for (;;) {
            int i = l, j = r, p = (l + r) / 2, ip = ix[p];
            double dp = d(ip);
            if (d(ix[l]) > dp) {
                ix[p] = ix[l];
                ix[l] = ip;
                ip = ix[p];
                dp = d(ip);
            }
            if (dp > d(ix[r])) {
                if (d(ix[l]) > d(ix[r])) {
                    ix[p] = ix[l];
                    ix[l] = ix[r];
                } else {
                    ix[p] = ix[r];
                }
                ix[r] = ip;
                ip = ix[p];
                dp = d(ip);
            }
            for (;;) {
                do ++i; while (d(ix[i]) < dp);
                do --j; while (d(ix[j]) > dp);
                if (i < j) std::swap(ix[i], ix[j]);
                else break;
            }
            if (r - j >= i - l && i - l > M) {
                ++istk;
                lstk[istk] = j + 1;
                rstk[istk] = r;
                r = i - 1;
            } else if (i - l > r - j && r - j > M) {
                ++istk;
                lstk[istk] = l;
                rstk[istk] = i - 1;
                l = j + 1;
            } else if (r - j > M) {
                l = j + 1;
            } else if (i - l > M) {
                r = i - 1;
            } else {
                if (istk < 1) break;
                l = lstk[istk];
                r = rstk[istk];
                --istk;
            }
        }
    }
    for (int i = 2; i <= n; ++i) {
        if (d(ix[i - 1]) > d(ix[i])) {
            int ip = ix[i], p = i - 1;
            double dp = d(ip);
            do {
                ix[p + 1] = ix[p];
                --p;
            } while (p > 0 && d(ix[p]) > dp);
            ix[p + 1] = ip;
        }
    }
    for (int i = 0; i < n; ++i) index[i] = ix[i + 1];
}

}  // namespace jdhip
