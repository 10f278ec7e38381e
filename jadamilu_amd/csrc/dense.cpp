// dense.cpp -- small dense linear algebra on the host (see dense.hpp).
#include "dense.hpp"

#include <cfloat>
#include <complex>

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

void gemv_host(int m, int nc, const double *a, int lda, const double *x, double *y, bool trans) {
    if (!trans) {
        for (int r = 0; r < m; ++r) y[r] = 0.0;
        for (int c = 0; c < nc; ++c)
            for (int r = 0; r < m; ++r) y[r] += a[r + (long)c * lda] * x[c];
    } else {
        for (int c = 0; c < nc; ++c) {
            double s = 0.0;
            for (int r = 0; r < m; ++r) s += a[r + (long)c * lda] * x[r];
            y[c] = s;
        }
    }
}

void ger_host(int m, int nc, double alpha, const double *x, const double *y, double *a,
              int lda) {
    for (int c = 0; c < nc; ++c) {
        const double t = alpha * y[c];
        for (int r = 0; r < m; ++r) a[r + (long)c * lda] += x[r] * t;
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

int chol_solve(int j, double *a, int lda, int k, double *b, int ldb) {
    for (int c = 0; c < j; ++c) {  // A = U^T U
        double d = a[c + c * lda];
        for (int q = 0; q < c; ++q) d -= a[q + c * lda] * a[q + c * lda];
        if (!(d > 0.0)) return c + 1;
        d = std::sqrt(d);
        a[c + c * lda] = d;
        for (int cc = c + 1; cc < j; ++cc) {
            double s = a[c + cc * lda];
            for (int q = 0; q < c; ++q) s -= a[q + c * lda] * a[q + cc * lda];
            a[c + cc * lda] = s / d;
        }
    }
    for (int r = 0; r < k; ++r) {
        double *x = b + (long)r * ldb;
        for (int i = 0; i < j; ++i) {  // U^T y = b
            double s = x[i];
            for (int q = 0; q < i; ++q) s -= a[q + i * lda] * x[q];
            x[i] = s / a[i + i * lda];
        }
        for (int i = j - 1; i >= 0; --i) {  // U x = y
            double s = x[i];
            for (int q = i + 1; q < j; ++q) s -= a[i + q * lda] * x[q];
            x[i] = s / a[i + i * lda];
        }
    }
    return 0;
}

namespace {

// H (n x n, column-major) to upper Hessenberg form by Householder
// reflections, H := P^T H P, accumulating Z := Z P.
void hessenberg(int n, std::vector<double> &H, std::vector<double> &Z) {
    auto h = [&](int i, int j) -> double & { return H[i + (size_t)j * n]; };
    auto z = [&](int i, int j) -> double & { return Z[i + (size_t)j * n]; };
    std::vector<double> v(n);
    for (int k = 0; k + 2 < n; ++k) {
        double xn2 = 0.0;
        for (int i = k + 2; i < n; ++i) xn2 += h(i, k) * h(i, k);
        if (xn2 == 0.0) continue;
        const double alpha = h(k + 1, k);
        const double nrm = std::sqrt(alpha * alpha + xn2);
        const double beta = (alpha >= 0.0) ? -nrm : nrm;
        v[k + 1] = alpha - beta;
        for (int i = k + 2; i < n; ++i) v[i] = h(i, k);
        const double tau = 2.0 / (v[k + 1] * v[k + 1] + xn2);
        for (int j = k; j < n; ++j) {
            double s = 0.0;
            for (int i = k + 1; i < n; ++i) s += v[i] * h(i, j);
            s *= tau;
            for (int i = k + 1; i < n; ++i) h(i, j) -= s * v[i];
        }
        for (int i = 0; i < n; ++i) {
            double s = 0.0, t = 0.0;
            for (int j = k + 1; j < n; ++j) {
                s += h(i, j) * v[j];
                t += z(i, j) * v[j];
            }
            s *= tau;
            t *= tau;
            for (int j = k + 1; j < n; ++j) {
                h(i, j) -= s * v[j];
                z(i, j) -= t * v[j];
            }
        }
        h(k + 1, k) = beta;
        for (int i = k + 2; i < n; ++i) h(i, k) = 0.0;
    }
}

// One reflector I - tau v v^T (v of length m = 2 or 3) applied to rows
// r0..r0+m-1 from the left over columns c0..n-1, and to the same columns
// from the right over rows 0..r1 of H and all rows of Z.
void apply_reflector(int n, std::vector<double> &H, std::vector<double> &Z, int m, int r0,
                     const double *v, double tau, int c0, int r1) {
    auto h = [&](int i, int j) -> double & { return H[i + (size_t)j * n]; };
    for (int j = c0; j < n; ++j) {
        double s = 0.0;
        for (int q = 0; q < m; ++q) s += v[q] * h(r0 + q, j);
        s *= tau;
        for (int q = 0; q < m; ++q) h(r0 + q, j) -= s * v[q];
    }
    for (int i = 0; i <= r1; ++i) {
        double s = 0.0;
        for (int q = 0; q < m; ++q) s += h(i, r0 + q) * v[q];
        s *= tau;
        for (int q = 0; q < m; ++q) h(i, r0 + q) -= s * v[q];
    }
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int q = 0; q < m; ++q) s += Z[i + (size_t)(r0 + q) * n] * v[q];
        s *= tau;
        for (int q = 0; q < m; ++q) Z[i + (size_t)(r0 + q) * n] -= s * v[q];
    }
}

// Householder vector of (x0, x1[, x2]) mapping it onto a multiple of e1;
// returns tau (0 for a zero vector).
double reflector(int m, const double *x, double *v) {
    double nrm2 = 0.0;
    for (int q = 0; q < m; ++q) nrm2 += x[q] * x[q];
    if (nrm2 == 0.0) return 0.0;
    const double nrm = std::sqrt(nrm2);
    const double beta = (x[0] >= 0.0) ? -nrm : nrm;
    v[0] = x[0] - beta;
    for (int q = 1; q < m; ++q) v[q] = x[q];
    double vtv = 0.0;
    for (int q = 0; q < m; ++q) vtv += v[q] * v[q];
    return 2.0 / vtv;
}

// Francis double-shift QR iteration on the Hessenberg H (Golub and Van Loan,
// Matrix Computations, 4th ed., algorithms 7.5.1-7.5.2), applied to the whole
// matrix so that H ends in real Schur form with Z the Schur vectors.
int francis_qr(int n, std::vector<double> &H, std::vector<double> &Z) {
    auto h = [&](int i, int j) -> double & { return H[i + (size_t)j * n]; };
    const double eps = DBL_EPSILON;
    double anorm = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= std::min(j + 1, n - 1); ++i) anorm += std::fabs(h(i, j));
    int hi = n - 1, its = 0, total = 0;
    while (hi >= 0) {
        int l = hi;
        for (; l > 0; --l) {
            double sc = std::fabs(h(l - 1, l - 1)) + std::fabs(h(l, l));
            if (sc == 0.0) sc = anorm;
            if (std::fabs(h(l, l - 1)) <= eps * sc) {
                h(l, l - 1) = 0.0;
                break;
            }
        }
        if (l == hi) {
            --hi;
            its = 0;
            continue;
        }
        if (l == hi - 1) {
            hi -= 2;
            its = 0;
            continue;
        }
        if (++total > 60 * n) return hi + 1;
        ++its;
        double s, t;
        if (its % 11 == 10) {  // exceptional shift
            const double w = std::fabs(h(hi, hi - 1)) + std::fabs(h(hi - 1, hi - 2));
            s = 1.5 * w;
            t = w * w;
        } else {
            const double a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1),
                         d = h(hi, hi);
            s = a + d;
            t = a * d - b * c;
        }
        double x[3];
        x[0] = h(l, l) * h(l, l) + h(l, l + 1) * h(l + 1, l) - s * h(l, l) + t;
        x[1] = h(l + 1, l) * (h(l, l) + h(l + 1, l + 1) - s);
        x[2] = h(l + 1, l) * h(l + 2, l + 1);
        double v[3];
        for (int k = l; k <= hi - 2; ++k) {
            const double tau = reflector(3, x, v);
            if (tau != 0.0)
                apply_reflector(n, H, Z, 3, k, v, tau, std::max(l, k - 1), std::min(k + 3, hi));
            x[0] = h(k + 1, k);
            x[1] = h(k + 2, k);
            if (k < hi - 2) x[2] = h(k + 3, k);
        }
        const double tau = reflector(2, x, v);
        if (tau != 0.0) apply_reflector(n, H, Z, 2, hi - 1, v, tau, hi - 2, hi);
        // the entries below the subdiagonal are zero in exact arithmetic
        for (int j = l; j < hi; ++j)
            for (int i = j + 2; i <= hi; ++i) h(i, j) = 0.0;
    }
    return 0;
}

}  // namespace

int real_eig(int j, double *a, int lda, double *wr, double *wi, double *v, int ldv) {
    using cd = std::complex<double>;
    const int n = j;
    if (n <= 0) return 0;
    std::vector<double> H((size_t)n * n), Z((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) H[r + (size_t)c * n] = a[r + (long)c * lda];
    for (int i = 0; i < n; ++i) Z[i + (size_t)i * n] = 1.0;
    hessenberg(n, H, Z);
    const int info = francis_qr(n, H, Z);
    if (info) return info;
    auto T = [&](int r, int c) { return H[r + (size_t)c * n]; };
    double tnorm = 0.0;
    for (int c = 0; c < n; ++c)
        for (int r = 0; r <= std::min(c + 1, n - 1); ++r) tnorm += std::fabs(T(r, c));
    const double smin = std::max(DBL_EPSILON * tnorm, DBL_MIN);
    // eigenvalues block by block
    std::vector<int> bstart(n);
    for (int i = 0; i < n;) {
        if (i + 1 < n && T(i + 1, i) != 0.0) {
            const double p = 0.5 * (T(i, i) - T(i + 1, i + 1));
            const double disc = p * p + T(i, i + 1) * T(i + 1, i);
            const double mid = 0.5 * (T(i, i) + T(i + 1, i + 1));
            if (disc >= 0.0) {
                const double r = std::sqrt(disc);
                wr[i] = mid + r;
                wr[i + 1] = mid - r;
                wi[i] = wi[i + 1] = 0.0;
            } else {
                wr[i] = wr[i + 1] = mid;
                wi[i] = std::sqrt(-disc);
                wi[i + 1] = -wi[i];
            }
            bstart[i] = bstart[i + 1] = i;
            i += 2;
        } else {
            wr[i] = T(i, i);
            wi[i] = 0.0;
            bstart[i] = i;
            ++i;
        }
    }
    std::vector<cd> y(n), x(n);
    for (int k = 0; k < n; ++k) {
        if (wi[k] < 0.0) continue;  // second of a complex pair: stored with the first
        const cd lam(wr[k], wi[k]);
        const int s = bstart[k];
        const bool two = (s + 1 < n && bstart[s + 1] == s);
        const int e = two ? s + 1 : s;
        std::fill(y.begin(), y.end(), cd(0.0));
        if (!two) {
            y[s] = 1.0;
        } else {
            const cd a0 = T(s, s) - lam, b0 = T(s, s + 1), c0 = T(s + 1, s),
                     d0 = T(s + 1, s + 1) - lam;
            if (std::abs(b0) + std::abs(a0) >= std::abs(c0) + std::abs(d0)) {
                y[s] = b0;
                y[s + 1] = -a0;
            } else {
                y[s] = -d0;
                y[s + 1] = c0;
            }
        }
        // back substitution over the blocks above
        for (int i = s - 1; i >= 0;) {
            if (i > 0 && bstart[i] == i - 1) {  // 2x2 block (i-1, i)
                cd r0 = 0.0, r1 = 0.0;
                for (int q = i + 1; q <= e; ++q) {
                    r0 -= T(i - 1, q) * y[q];
                    r1 -= T(i, q) * y[q];
                }
                const cd m00 = T(i - 1, i - 1) - lam, m01 = T(i - 1, i), m10 = T(i, i - 1),
                         m11 = T(i, i) - lam;
                cd det = m00 * m11 - m01 * m10;
                if (std::abs(det) < smin) det = smin;
                y[i - 1] = (r0 * m11 - m01 * r1) / det;
                y[i] = (m00 * r1 - m10 * r0) / det;
                i -= 2;
            } else {
                cd r = 0.0;
                for (int q = i + 1; q <= e; ++q) r -= T(i, q) * y[q];
                cd den = T(i, i) - lam;
                if (std::abs(den) < smin) den = smin;
                y[i] = r / den;
                --i;
            }
        }
        for (int r = 0; r < n; ++r) {
            cd t = 0.0;
            for (int q = 0; q <= e; ++q) t += Z[r + (size_t)q * n] * y[q];
            x[r] = t;
        }
        double nrm = 0.0;
        for (int r = 0; r < n; ++r) nrm += std::norm(x[r]);
        nrm = std::sqrt(nrm);
        if (wi[k] == 0.0) {
            for (int r = 0; r < n; ++r) v[r + (long)k * ldv] = x[r].real() / nrm;
        } else {
            int m = 0;
            for (int r = 1; r < n; ++r)
                if (std::norm(x[r]) > std::norm(x[m])) m = r;
            const cd rot = std::conj(x[m]) / (std::abs(x[m]) * nrm);
            for (int r = 0; r < n; ++r) {
                const cd t = x[r] * rot;
                v[r + (long)k * ldv] = t.real();
                v[r + (long)(k + 1) * ldv] = t.imag();
            }
        }
    }
    (void)a;
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
