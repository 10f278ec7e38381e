/* jdo_internal.h -- private declarations of the CPU oracle (test
 * infrastructure only, see jd_oracle.h). */
#ifndef JDO_INTERNAL_H
#define JDO_INTERNAL_H

double jdo_dnrm2(int n, const double *x, int incx);
double jdo_ddot(int n, const double *x, const double *y);
void jdo_daxpy(int n, double a, const double *x, double *y);
void jdo_dscal(int n, double a, double *x);
void jdo_gemv_t(int n, int jv, const double *v, long ldv, const double *w,
                double *y);
void jdo_gemv_n(int n, int jv, double alpha, const double *v, long ldv,
                const double *y, double beta, double *w);
void jdo_syr2k_ut(int j, int n, const double *a, long lda, const double *b,
                  long ldb, double *c, int ldc);
void jdo_syrk_ut(int j, int n, const double *a, long lda, double *c, int ldc);
void jdo_symv_u(int j, const double *a, int lda, const double *x, double *y);
void jdo_syr2_u(int j, double alpha, const double *x, const double *y,
                double *a, int lda);
void jdo_geqrf(int m, int k, double *a, int lda, double *tau);
void jdo_ormqr_rn(int nr, int k, int nref, const double *a, int lda,
                  const double *tau, double *c, long ldc);
void jdo_ormqr_lt(int k, int nc, int nref, const double *a, int lda,
                  const double *tau, double *c, int ldc);

#endif
