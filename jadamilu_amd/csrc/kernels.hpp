// kernels.hpp -- launch wrappers for the gfx950 kernels of the JD hot path.
//
// All kernels run on the caller's HIP stream.  Reductions are deterministic:
// every block writes its partial sums, the last block to finish (ticket from
// a device counter) sums the partials in a fixed order and writes the result
// into device memory, so a dependent kernel can consume it without a host
// round trip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace jdhip {

constexpr int kBlock = 256;      // threads per block (4 wave64)
constexpr int kMaxCols = 64;     // max columns of a tall-skinny product
constexpr int kMaxRedBlocks = 1024;

// Device-resident scalars of the inner solver (one per engine).
struct InnerScalars {
    double rho, rho1, alpha, pq, alprho, gk2, xn2, eta;
    double ggk2, ttk2, ttk2m, ck2;
    int iter;    // unused on device (the host passes the iteration number)
    int posdef;  // 1: PCG, 0: sQMR
    int flag;    // 0 ok, 1: rho breakdown (before matvec), 2: alpha breakdown
    int pad;
};

// Reduction scratch owned by an engine.
struct RedBuf {
    double *partials;   // kMaxRedBlocks * kMaxCols
    unsigned *counter;  // 8 counters
    double *out;        // kMaxCols results
};

int red_blocks(long n);

// ---- BLAS-1 / tall-skinny building blocks (double) ----
void fill(hipStream_t s, double *x, double v, long n);
void scal(hipStream_t s, double *x, double a, long n);
void axpy(hipStream_t s, double *y, double a, const double *x, long n);
// out[c] = sum_i V[c*ldv+i] * w[i], c < jv  (result in rb.out, device)
void multidot(hipStream_t s, const RedBuf &rb, const double *V, long ldv, int jv,
              const double *w, long n, double *out);
// w = beta*w + alpha * V * y (y device, jv entries); if nrm2out, also
// nrm2out[0] = ||w||^2 after the update.
void gemv_n(hipStream_t s, const RedBuf &rb, double *w, const double *V, long ldv,
            int jv, const double *y, double alpha, double beta, long n,
            double *nrm2out);
// y += a x ; out[0] = ||y||^2
void axpy_nrm2(hipStream_t s, const RedBuf &rb, double *y, double a, const double *x, long n,
               double *out);
// out[0] = x.y
void dot(hipStream_t s, const RedBuf &rb, const double *x, const double *y, long n,
         double *out);
// out[j*ldo+i] = sum_k V[k*ldv+i] Q[k+j*ldq]   (n x K times K x J)
void tall_gemm(hipStream_t s, const double *V, long ldv, int K, const double *Q,
               int ldq, int J, double *out, long ldo, long n);
// out = {index of max |x_i| (first), x_idx} as doubles
void argmax_abs(hipStream_t s, const RedBuf &rb, const double *x, long n, double *out);

// ---- sparse ----
struct CsrDev {
    int n = 0;
    long nnz = 0;
    int *rowptr = nullptr;  // 0-based, n+1
    int *col = nullptr;
    double *val = nullptr;
    float *valf = nullptr;  // optional FP32 copy
    int lpr = 8;            // lanes per row
};
// y = A x ; if eta_ptr: y -= (*eta_ptr) x ; if u: out[0] = u.y (after)
void spmv(hipStream_t s, const RedBuf &rb, const CsrDev &A, const double *x, double *y,
          const double *eta_dev, double eta_host, const double *u, double *udot_out);

// ---- preconditioners ----
void diag_apply(hipStream_t s, double *z, const double *d, const double *r, long n);

// ---- inner solver (LSOLREVCOM) fused steps ----
// z = P^-1 r is already in z.  y = Q^T z (jq columns) -> rb.out region yout
// then z -= Q y ; rho = r.z ; checks ; p = z (+ beta p)
void cg_project_rho(hipStream_t s, const RedBuf &rb, double *z, const double *Q, int jq,
                    const double *r, long n, InnerScalars *sc, double *ycoef);
void cg_direction(hipStream_t s, double *z, double *p, long n, InnerScalars *sc, int iter);
// q already = (A - eta) p ; alprho = u.q (device, computed by spmv) ;
// q -= u alprho ; pq = p.q ; alpha = rho/pq with checks
void cg_project_alpha(hipStream_t s, const RedBuf &rb, double *q, const double *u,
                      const double *p, long n, InnerScalars *sc);
// PCG: r -= alpha q ; gk2 = r.r ; x += alpha p ; xn2 = x.x
void cg_update_posdef(hipStream_t s, const RedBuf &rb, double *r, const double *q,
                      double *x, const double *p, long n, InnerScalars *sc);
// sQMR: r -= alpha q ; gk2 = r.r, then the smoothing update of x, x2
void qmr_update_r(hipStream_t s, const RedBuf &rb, double *r, const double *q, long n,
                  InnerScalars *sc);
void qmr_update_x(hipStream_t s, const RedBuf &rb, double *x, double *x2, const double *p,
                  long n, InnerScalars *sc, int iter);

}  // namespace jdhip
