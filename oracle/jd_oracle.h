/*
 * jd_oracle.h -- CPU oracle for the Jacobi-Davidson hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, single-threaded restatement
 * of the reference algorithm (src/jd/JDaux.F, src/jd/PJD.F,
 * src/jd/dglprec*.c).  It is used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg as the checker, and never by the product path.
 *
 * The calling sequence mirrors the reference's public drivers:
 *   jdo_pjdrevcom  <-> DPJDREVCOM  (PJD.F:268)
 *   jdo_jdrevcom   <-> DJDREVCOM   (PJD.F:369)
 *   jdo_pjd        <-> DPJD        (PJD.F:5), diagonal/ILDL preconditioner
 *   jdo_cleanup    <-> DPJDCLEANUP (PJD.F:435)
 * All integer arguments are int32 (the reference's INT32/INT64N platforms),
 * reals are double, matrix input is 1-based upper-triangle CSR.
 */
#ifndef JD_ORACLE_H
#define JD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct jdo_state jdo_state;

jdo_state *jdo_create(void);
void jdo_destroy(jdo_state *s);

/* DPJDREVCOM (PJD.F:268-362).  ja[0] < 0 selects the diagonal preconditioner
 * with a[0..n-1] the diagonal.  Returns with *ijob==1 when
 * X(ndx2:ndx2+n-1) = A*X(ndx1:ndx1+n-1) is needed (1-based NDX). */
void jdo_pjdrevcom(jdo_state *s, int n, double *a, int *ja, int *ia,
                   double *eigs, double *res, double *x, int lx, int *neig,
                   double sigma, int isearch, int ninit, int madspace,
                   int *iter, double tol, double *shift, double *droptol,
                   double *mem, int *icntl, int *ijob, int *ndx1, int *ndx2,
                   int iprint, int *info, double *gap);

/* DJDREVCOM (PJD.F:369-427): matvec (ijob=1) and preconditioner solve
 * (ijob=2: X(ndx1) = P^-1 X(ndx2)) both by reverse communication. */
void jdo_jdrevcom(jdo_state *s, int n, double *eigs, double *res, double *x,
                  int lx, int *neig, double sigma, int isearch, int ninit,
                  int madspace, int *iter, double tol, int *ijob, int *ndx1,
                  int *ndx2, int iprint, int *info, double *gap);

/* DPJD (PJD.F:5-259) with the matvec done internally by MATVECA
 * (JDaux.F:3397) on the upper-triangle CSR (a, ja, ia), diagonal first in
 * each row after the call (the reference's row permutation). */
void jdo_pjd(jdo_state *s, int n, double *a, int *ja, int *ia, double *eigs,
             double *res, double *x, int lx, int *neig, double sigma,
             int isearch, int ninit, int madspace, int *iter, double tol,
             double *shift, double *droptol, double *mem, int *icntl,
             int iprint, int *info, double *gap);

void jdo_cleanup(jdo_state *s);

/* MATVECA (JDaux.F:3397-3418): y = A x, symmetric A stored as 1-based upper
 * CSR with the diagonal in the first slot of every row. */
void jdo_matveca(int n, const int *ia, const int *ja, const double *a,
                 const double *x, double *y);

/* Generic symmetric upper-CSR matvec (diagonal anywhere in the row), as the
 * user routine of samples/EXAMPLE9.f does it. */
void jdo_matvec_any(int n, const int *ia, const int *ja, const double *a,
                    const double *x, double *y);

/* Diagnostics recorded by the oracle: every preconditioner (re)setup shift
 * (the "used shift:" lines of PJDINIT, JDaux.F:713-717), and the number of
 * inner (PCG/sQMR) iterations. */
int jdo_nshift(const jdo_state *s);
double jdo_shift(const jdo_state *s, int i);
long jdo_inner_iters(const jdo_state *s);
int jdo_nprecond(const jdo_state *s);

/* Dense helpers exported for their own tests. */
int jdo_syev(int j, double *a, int lda, double *w);
int jdo_sygv(int j, double *a, int lda, double *b, int ldb, double *w);
void jdo_sortrx(int n, const double *data, int *index);

/* A user preconditioner installed for jdo_pjd-style internal solves
 * (used by tests that drive the ILDL apply through the oracle). */
typedef void (*jdo_psolve_fn)(void *ctx, int n, double *out, const double *in);
void jdo_set_psolve(jdo_state *s, jdo_psolve_fn fn, void *ctx);

#ifdef __cplusplus
}
#endif
#endif
