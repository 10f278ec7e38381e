/*
 * jdo_dense.c -- dense BLAS/LAPACK-equivalents used by the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see jd_oracle.h).
 *
 * The reference calls netlib BLAS/LAPACK (macro list in
 * src/include/ilupackmacros.h:55-72).  This file restates, in plain C, the
 * few routines the JD engine needs, with the operation order of the
 * reference-BLAS loops where that order is observable:
 *   dnrm2   scaled sum of squares (Hammarling's algorithm)
 *   dgemv   'T' as per-column dot, 'N' as column axpys
 *   dsyr2k/dsyrk/dsymv/dsyr2   upper triangle only
 *   dgeqrf/dormqr  unblocked Householder (dgeqr2/dorm2r): LAPACK takes the
 *                  unblocked path whenever the block size (32) exceeds the
 *                  number of reflectors, which always holds here (J <= 25)
 *   dsyev/dsygv    replaced by cyclic Jacobi (eigenvalues agree with LAPACK
 *                  to rounding; eigenvectors agree up to sign, and the JD
 *                  iteration is invariant under eigenvector signs)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "jdo_internal.h"

double jdo_dnrm2(int n, const double *x, int incx)
{
    double scale = 0.0, ssq = 1.0;
    if (n < 1) return 0.0;
    if (n == 1) return fabs(x[0]);
    for (int i = 0; i < n; i++) {
        double xi = x[(long)i * incx];
        if (xi != 0.0) {
            double ax = fabs(xi);
            if (scale < ax) {
                double r = scale / ax;
                ssq = 1.0 + ssq * r * r;
                scale = ax;
            } else {
                double r = ax / scale;
                ssq += r * r;
            }
        }
    }
    return scale * sqrt(ssq);
}

double jdo_ddot(int n, const double *x, const double *y)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

void jdo_daxpy(int n, double a, const double *x, double *y)
{
    if (a == 0.0) return;
    for (int i = 0; i < n; i++) y[i] += a * x[i];
}

void jdo_dscal(int n, double a, double *x)
{
    for (int i = 0; i < n; i++) x[i] *= a;
}

/* y(0:jv) = V^T w (dgemv 'T', alpha 1, beta 0) */
void jdo_gemv_t(int n, int jv, const double *v, long ldv, const double *w,
                double *y)
{
    for (int j = 0; j < jv; j++) {
        const double *c = v + (long)j * ldv;
        double t = 0.0;
        for (int i = 0; i < n; i++) t += c[i] * w[i];
        y[j] = t;
    }
}

/* w = beta*w + alpha*V*y (dgemv 'N') */
void jdo_gemv_n(int n, int jv, double alpha, const double *v, long ldv,
                const double *y, double beta, double *w)
{
    if (beta == 0.0) {
        for (int i = 0; i < n; i++) w[i] = 0.0;
    } else if (beta != 1.0) {
        for (int i = 0; i < n; i++) w[i] *= beta;
    }
    for (int j = 0; j < jv; j++) {
        if (y[j] != 0.0) {
            double t = alpha * y[j];
            const double *c = v + (long)j * ldv;
            for (int i = 0; i < n; i++) w[i] += t * c[i];
        }
    }
}

/* C(0:j,0:j) upper = A^T B + B^T A  (dsyr2k 'U','T', alpha 1, beta 0) */
void jdo_syr2k_ut(int j, int n, const double *a, long lda, const double *b,
                  long ldb, double *c, int ldc)
{
    for (int jj = 0; jj < j; jj++)
        for (int i = 0; i <= jj; i++) {
            double t1 = 0.0, t2 = 0.0;
            const double *ai = a + (long)i * lda, *bj = b + (long)jj * ldb;
            const double *bi = b + (long)i * ldb, *aj = a + (long)jj * lda;
            for (int l = 0; l < n; l++) {
                t1 += ai[l] * bj[l];
                t2 += bi[l] * aj[l];
            }
            c[i + jj * ldc] = t1 + t2;
        }
}

/* C upper = A^T A (dsyrk 'U','T', alpha 1, beta 0) */
void jdo_syrk_ut(int j, int n, const double *a, long lda, double *c, int ldc)
{
    for (int jj = 0; jj < j; jj++)
        for (int i = 0; i <= jj; i++) {
            double t = 0.0;
            const double *ai = a + (long)i * lda, *aj = a + (long)jj * lda;
            for (int l = 0; l < n; l++) t += ai[l] * aj[l];
            c[i + jj * ldc] = t;
        }
}

/* y = A x with A symmetric, upper triangle referenced (dsymv 'U', alpha 1,
 * beta 0) */
void jdo_symv_u(int j, const double *a, int lda, const double *x, double *y)
{
    for (int i = 0; i < j; i++) y[i] = 0.0;
    for (int jj = 0; jj < j; jj++) {
        double t1 = x[jj], t2 = 0.0;
        for (int i = 0; i < jj; i++) {
            y[i] += t1 * a[i + jj * lda];
            t2 += a[i + jj * lda] * x[i];
        }
        y[jj] += t1 * a[jj + jj * lda] + t2;
    }
}

/* A upper += alpha*(x y^T + y x^T) (dsyr2 'U') */
void jdo_syr2_u(int j, double alpha, const double *x, const double *y,
                double *a, int lda)
{
    for (int jj = 0; jj < j; jj++) {
        if (x[jj] != 0.0 || y[jj] != 0.0) {
            double t1 = alpha * y[jj], t2 = alpha * x[jj];
            for (int i = 0; i <= jj; i++)
                a[i + jj * lda] += x[i] * t1 + y[i] * t2;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Symmetric eigenproblem by cyclic Jacobi rotations.                      */

static int jacobi_full(int j, double *m, double *v, double *w)
{
    for (int c = 0; c < j; c++)
        for (int r = 0; r < j; r++) v[r + c * j] = (r == c) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, dg = 0.0;
        for (int q = 0; q < j; q++) {
            dg += m[q + q * j] * m[q + q * j];
            for (int p = 0; p < q; p++) off += m[p + q * j] * m[p + q * j];
        }
        if (off == 0.0 || off <= 1e-36 * dg) break;
        for (int p = 0; p < j - 1; p++)
            for (int q = p + 1; q < j; q++) {
                double apq = m[p + q * j];
                if (apq == 0.0) continue;
                double app = m[p + p * j], aqq = m[q + q * j];
                double g = 100.0 * fabs(apq);
                if (sweep > 4 && fabs(app) + g == fabs(app) &&
                    fabs(aqq) + g == fabs(aqq)) {
                    m[p + q * j] = m[q + p * j] = 0.0;
                    continue;
                }
                double th = (aqq - app) / (2.0 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) /
                           (fabs(th) + sqrt(th * th + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                double tau = sn / (1.0 + cs);
                m[p + p * j] = app - t * apq;
                m[q + q * j] = aqq + t * apq;
                m[p + q * j] = m[q + p * j] = 0.0;
                for (int r = 0; r < j; r++) {
                    if (r == p || r == q) continue;
                    double arp = m[r + p * j], arq = m[r + q * j];
                    double nrp = arp - sn * (arq + tau * arp);
                    double nrq = arq + sn * (arp - tau * arq);
                    m[r + p * j] = m[p + r * j] = nrp;
                    m[r + q * j] = m[q + r * j] = nrq;
                }
                for (int r = 0; r < j; r++) {
                    double vrp = v[r + p * j], vrq = v[r + q * j];
                    v[r + p * j] = vrp - sn * (vrq + tau * vrp);
                    v[r + q * j] = vrq + sn * (vrp - tau * vrq);
                }
            }
    }
    for (int i = 0; i < j; i++) w[i] = m[i + i * j];
    /* ascending selection sort, ties keep their order */
    for (int i = 0; i < j - 1; i++) {
        int k = i;
        for (int l = i + 1; l < j; l++)
            if (w[l] < w[k]) k = l;
        if (k != i) {
            double t = w[i];
            w[i] = w[k];
            w[k] = t;
            for (int r = 0; r < j; r++) {
                t = v[r + i * j];
                v[r + i * j] = v[r + k * j];
                v[r + k * j] = t;
            }
        }
    }
    return 0;
}

/* dsyev('V','U'): eigenvalues ascending in w, eigenvectors over a. */
int jdo_syev(int j, double *a, int lda, double *w)
{
    double *m = (double *)malloc(sizeof(double) * j * j * 2);
    double *v = m + j * j;
    for (int c = 0; c < j; c++)
        for (int r = 0; r <= c; r++) m[r + c * j] = m[c + r * j] = a[r + c * lda];
    jacobi_full(j, m, v, w);
    for (int c = 0; c < j; c++)
        for (int r = 0; r < j; r++) a[r + c * lda] = v[r + c * j];
    free(m);
    return 0;
}

/* dsygv(itype 1,'V','U'): A y = lambda B y, B = U^T U (Cholesky, upper),
 * eigenvalues ascending, eigenvectors B-normalised over a; b holds U on
 * exit.  Returns j+k if the leading minor of order k of B is not positive
 * definite (LAPACK convention). */
int jdo_sygv(int j, double *a, int lda, double *b, int ldb, double *w)
{
    for (int c = 0; c < j; c++) {
        double d = b[c + c * ldb];
        for (int k = 0; k < c; k++) d -= b[k + c * ldb] * b[k + c * ldb];
        if (!(d > 0.0)) return j + c + 1;
        d = sqrt(d);
        b[c + c * ldb] = d;
        for (int cc = c + 1; cc < j; cc++) {
            double s = b[c + cc * ldb];
            for (int k = 0; k < c; k++) s -= b[k + c * ldb] * b[k + cc * ldb];
            b[c + cc * ldb] = s / d;
        }
    }
    double *m = (double *)malloc(sizeof(double) * j * j * 3);
    double *v = m + j * j, *t = v + j * j;
    for (int c = 0; c < j; c++)
        for (int r = 0; r <= c; r++) m[r + c * j] = m[c + r * j] = a[r + c * lda];
    /* t = U^-T m  (forward substitution with the lower factor U^T) */
    for (int c = 0; c < j; c++)
        for (int r = 0; r < j; r++) {
            double s = m[r + c * j];
            for (int k = 0; k < r; k++) s -= b[k + r * ldb] * t[k + c * j];
            t[r + c * j] = s / b[r + r * ldb];
        }
    /* m = t U^-1, i.e. m^T = U^-T t^T */
    for (int r = 0; r < j; r++)
        for (int c = 0; c < j; c++) {
            double s = t[r + c * j];
            for (int k = 0; k < c; k++) s -= b[k + c * ldb] * m[r + k * j];
            m[r + c * j] = s / b[c + c * ldb];
        }
    for (int c = 0; c < j; c++)
        for (int r = 0; r < c; r++) {
            double s = 0.5 * (m[r + c * j] + m[c + r * j]);
            m[r + c * j] = m[c + r * j] = s;
        }
    jacobi_full(j, m, v, w);
    /* y = U^-1 z (back substitution) */
    for (int c = 0; c < j; c++)
        for (int r = j - 1; r >= 0; r--) {
            double s = v[r + c * j];
            for (int k = r + 1; k < j; k++) s -= b[r + k * ldb] * a[k + c * lda];
            a[r + c * lda] = s / b[r + r * ldb];
        }
    free(m);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Householder QR (dgeqr2) and application of Q (dorm2r).                  */

static void larfg(int n, double *alpha, double *x, double *tau)
{
    if (n <= 1) {
        *tau = 0.0;
        return;
    }
    double xn = jdo_dnrm2(n - 1, x, 1);
    if (xn == 0.0) {
        *tau = 0.0;
        return;
    }
    double beta = -copysign(hypot(*alpha, xn), *alpha);
    *tau = (beta - *alpha) / beta;
    jdo_dscal(n - 1, 1.0 / (*alpha - beta), x);
    *alpha = beta;
}

/* QR of the m x k block a (column-major, lda); reflectors below the
 * diagonal, tau[0:k]. */
void jdo_geqrf(int m, int k, double *a, int lda, double *tau)
{
    int kk = m < k ? m : k;
    for (int i = 0; i < kk; i++) {
        double *col = a + i + (long)i * lda;
        larfg(m - i, col, col + 1, &tau[i]);
        if (i < k - 1) {
            double aii = col[0];
            col[0] = 1.0;
            for (int c = i + 1; c < k; c++) {
                double *cc = a + i + (long)c * lda;
                double w = 0.0;
                for (int r = 0; r < m - i; r++) w += cc[r] * col[r];
                for (int r = 0; r < m - i; r++) cc[r] -= tau[i] * w * col[r];
            }
            col[0] = aii;
        }
    }
}

/* C (nr x k) := C * Q, Q = H_1 ... H_nref from jdo_geqrf of a k x nref
 * block (dorm2r 'R','N'). */
void jdo_ormqr_rn(int nr, int k, int nref, const double *a, int lda,
                  const double *tau, double *c, long ldc)
{
    double *v = (double *)malloc(sizeof(double) * k);
    double *w = (double *)malloc(sizeof(double) * (nr > 0 ? nr : 1));
    for (int i = 0; i < nref; i++) {
        if (tau[i] == 0.0) continue;
        int len = k - i;
        v[0] = 1.0;
        for (int r = 1; r < len; r++) v[r] = a[i + r + (long)i * lda];
        /* w = C(:, i:k) v  (column by column, like dgemv 'N') */
        for (int r = 0; r < nr; r++) w[r] = 0.0;
        for (int cc = 0; cc < len; cc++) {
            double t = v[cc];
            const double *col = c + (long)(i + cc) * ldc;
            for (int r = 0; r < nr; r++) w[r] += t * col[r];
        }
        /* C(:, i:k) -= tau w v^T  (dger) */
        for (int cc = 0; cc < len; cc++) {
            double t = -tau[i] * v[cc];
            double *col = c + (long)(i + cc) * ldc;
            for (int r = 0; r < nr; r++) col[r] += w[r] * t;
        }
    }
    free(v);
    free(w);
}

/* C (k x nc) := Q^T C (dorm2r 'L','T'). */
void jdo_ormqr_lt(int k, int nc, int nref, const double *a, int lda,
                  const double *tau, double *c, int ldc)
{
    double *v = (double *)malloc(sizeof(double) * k);
    for (int i = 0; i < nref; i++) {
        if (tau[i] == 0.0) continue;
        int len = k - i;
        v[0] = 1.0;
        for (int r = 1; r < len; r++) v[r] = a[i + r + (long)i * lda];
        for (int cc = 0; cc < nc; cc++) {
            double *col = c + i + (long)cc * ldc;
            double w = 0.0;
            for (int r = 0; r < len; r++) w += col[r] * v[r];
            w *= -tau[i];
            for (int r = 0; r < len; r++) col[r] += v[r] * w;
        }
    }
    free(v);
}

/* ---------------------------------------------------------------------- */
/* SORTRX (JDaux.F:2762-2987): index sort, median-of-three quicksort with   */
/* insertion-sort finish (Knuth Vol. 3, Algorithm Q), cut-off 9.           */
void jdo_sortrx(int n, const double *data, int *index)
{
    /* 1-based restatement; data[] and index[] are used with an offset. */
    const int M = 9;
    int lstk[64], rstk[64], istk = 0;
    int *ix = index - 1;
    const double *d = data - 1;
    for (int i = 1; i <= n; i++) ix[i] = i;
    if (n > M) {
        int l = 1, r = n;
        for (;;) {
            int i = l, j = r;
            int p = (l + r) / 2;
            int ip = ix[p];
            double dp = d[ip];
            if (d[ix[l]] > dp) {
                ix[p] = ix[l];
                ix[l] = ip;
                ip = ix[p];
                dp = d[ip];
            }
            if (dp > d[ix[r]]) {
                if (d[ix[l]] > d[ix[r]]) {
                    ix[p] = ix[l];
                    ix[l] = ix[r];
                } else {
                    ix[p] = ix[r];
                }
                ix[r] = ip;
                ip = ix[p];
                dp = d[ip];
            }
            for (;;) {
                do { i++; } while (d[ix[i]] < dp);
                do { j--; } while (d[ix[j]] > dp);
                if (i < j) {
                    int t = ix[i];
                    ix[i] = ix[j];
                    ix[j] = t;
                } else {
                    break;
                }
            }
            if (r - j >= i - l && i - l > M) {
                istk++;
                lstk[istk] = j + 1;
                rstk[istk] = r;
                r = i - 1;
            } else if (i - l > r - j && r - j > M) {
                istk++;
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
                istk--;
            }
        }
    }
    for (int i = 2; i <= n; i++) {
        if (d[ix[i - 1]] > d[ix[i]]) {
            int ip = ix[i];
            double dp = d[ip];
            int p = i - 1;
            for (;;) {
                ix[p + 1] = ix[p];
                p--;
                if (p > 0 && d[ix[p]] > dp) continue;
                break;
            }
            ix[p + 1] = ip;
        }
    }
    for (int i = 0; i < n; i++) index[i] = ix[i + 1];
}
