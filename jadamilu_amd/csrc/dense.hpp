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
// A(upper) += alpha (x y^T + y x^T)  (dsyr2 'U')
void syr2_upper(int j, double alpha, const double *x, const double *y, double *a, int lda);
// Symmetric eigenproblem (upper triangle of a read): ascending w, vectors
// over a (cyclic Jacobi; replaces dsyev).
int sym_eig(int j, double *a, int lda, double *w);
// Generalised problem A y = lambda B y, B SPD (replaces dsygv itype 1):
// ascending w, B-normalised vectors over a, Cholesky factor over b.  Returns
// j+k when the leading minor k of B is not positive definite.
int sym_gen_eig(int j, double *a, int lda, double *b, int ldb, double *w);
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
