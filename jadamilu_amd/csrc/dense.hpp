// dense.hpp -- small dense linear algebra on the host (J x J, J <= MADSPACE).
//
// The projected problems of the JD iteration are at most MADSPACE x MADSPACE
// (25 at the benchmark settings): they stay on the host, where a full solve
// costs microseconds, and only n-length data lives on the GPU.
#pragma once
#include <vector>

namespace jdhip {

double small_nrm2(int n, const double *x, int inc = 1);  // scaled (dnrm2)
double small_dot(int n, const double *x, const double *y);
// y = A x, A symmetric with the upper triangle referenced (dsymv 'U')
void symv_upper(int j, const double *a, int lda, const double *x, double *y);
// y = A x (trans = false) or y = A^T x, A m x n general (dgemv)
void gemv_host(int m, int nc, const double *a, int lda, const double *x, double *y, bool trans);
// A += alpha x y^T, A m x n (dger)
void ger_host(int m, int nc, double alpha, const double *x, const double *y, double *a,
              int lda);
// A(upper) += alpha (x y^T + y x^T)  (dsyr2 'U')
void syr2_upper(int j, double alpha, const double *x, const double *y, double *a, int lda);
// Symmetric eigenproblem (upper triangle of a read): ascending w, vectors
// over a (cyclic Jacobi; replaces dsyev).
int sym_eig(int j, double *a, int lda, double *w);
// Generalised problem A y = lambda B y, B SPD (replaces dsygv itype 1):
// ascending w, B-normalised vectors over a, Cholesky factor over b.  Returns
// j+k when the leading minor k of B is not positive definite.
int sym_gen_eig(int j, double *a, int lda, double *b, int ldb, double *w);
// A X = B for SPD A (upper triangle read, Cholesky factor left in it) and k
// right-hand sides over b (replaces dposv 'U').  Returns c > 0 when the
// leading minor c of A is not positive definite.
int chol_solve(int j, double *a, int lda, int k, double *b, int ldb);
// Real general eigenproblem (replaces dgeev with JOBVL = 'N', JOBVR = 'V'):
// Householder reduction to Hessenberg form, Francis double-shift QR to the
// real Schur form with accumulated Schur vectors, eigenvectors by back
// substitution.  wr + i wi in Schur order; v (ld ldv): a real eigenvalue's
// vector in its column, a complex pair (wi > 0 first) as the real and
// imaginary parts of the first one's vector in two columns; unit Euclidean
// norm, a complex vector's largest component real.  a is overwritten.
// Returns 0, or k > 0 when the QR iteration did not converge.
int real_eig(int j, double *a, int lda, double *wr, double *wi, double *v, int ldv);
// Householder QR of the m x k block a (reflectors below the diagonal).
void house_qr(int m, int k, double *a, int lda, double *tau);
// C (nr x m) := C Q  with Q = H_1..H_nref of an m-row QR.
void apply_q_right(int nr, int m, int nref, const double *a, int lda, const double *tau,
                   double *c, int ldc);
// C (m x nc) := Q^T C
void apply_qt_left(int m, int nc, int nref, const double *a, int lda, const double *tau,
                   double *c, int ldc);
// Explicit first `ncol` columns of Q (m x ncol, ld m).
void form_q(int m, int ncol, int nref, const double *a, int lda, const double *tau,
            double *q);
// SORTRX: index sort (1-based indices), median-of-three quicksort with an
// insertion-sort finish, as src/jd/JDaux.F:2762-2987.
void sortrx(int n, const double *data, int *index);

}  // namespace jdhip
