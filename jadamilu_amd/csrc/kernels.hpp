// kernels.hpp -- launch wrappers for the gfx950 kernels of the JD hot path.
//
// All kernels run on the caller's HIP stream.  Reductions are deterministic:
// every block writes its partial sums, the last block to finish (ticket from
// a device counter) sums the partials in a fixed order and writes the result
// into device memory, so a dependent kernel can consume it without a host
// round trip.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>

namespace jdhip {

constexpr int kBlock = 256;      // threads per block (4 wave64)
constexpr int kMaxCols = 64;     // max columns of a tall-skinny product
constexpr int kMaxRedBlocks = 4096;  // enough waves in flight to cover HBM latency
constexpr int kMaxPartials = kMaxRedBlocks * 64;  // reduction scratch (doubles)

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
    double *partials;   // kMaxPartials
    unsigned *counter;  // 8 counters
    double *out;        // kMaxCols results
};

int red_blocks(long n);

// ---- BLAS-1 / tall-skinny building blocks (double) ----
void fill(hipStream_t s, double *x, double v, long n);
void scal(hipStream_t s, double *x, double a, long n);
void axpy(hipStream_t s, double *y, double a, const double *x, long n);
// out[c] = sum_i V[c*ldv+i] * w[i], c < jv (device, accumulated in double)
template <typename T>
void multidot(hipStream_t s, const RedBuf &rb, const T *V, long ldv, int jv, const T *w, long n,
              double *out);
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
// ---- FP64 matrix-core (MFMA) contractions, kernels_mfma.hip ----
// Partial sums of gram() hold at most this many doubles (4 MiB).
constexpr long kGramScratch = 524288;
// Workgroups of gram() for these sizes: at most 1024, at most one per 64 rows,
// and nblk * K * M <= kGramScratch.
int gram_blocks(long n, int K, int M);
// C(i + j*ldc) = sum_r V[i*ldv + r] W[j*ldw + r]  (C = V^T W, K, M <= 64),
// with gram_blocks(n, K, M) * K * M doubles of partial sums in scratch
// (device).  Deterministic (the same summation order on every call).  0 on success,
// -1 if K or M exceeds 64.
int gram(hipStream_t s, long n, int K, int M, const double *V, long ldv, const double *W,
         long ldw, double *C, int ldc, double *scratch);
// Y = V Q on the matrix cores (K, J <= 64); same result layout as tall_gemm.
int tall_gemm_mfma(hipStream_t s, const double *V, long ldv, int K, const double *Q, int ldq,
                   int J, double *Y, long ldy, long n);
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
    int lpr = 8;            // lanes per row (CSR kernel); 0: sliced-ELL kernel
    // Sliced ELL (slices of 64 consecutive rows, one wave64 per slice, one
    // lane per row, entries of a slice stored column-major so every load of
    // the wave is contiguous).  Column indices are 16-bit offsets from the
    // row when the whole slice allows it, otherwise 32-bit columns.
    int nslices = 0;
    long *s_vptr = nullptr;   // element offset of each slice's values
    long *s_iptr = nullptr;   // byte offset of each slice's indices
    int *s_wm = nullptr;      // 2*width + (1 if 16-bit offsets)
    double *s_val = nullptr;
    float *s_valf = nullptr;
    char *s_idx = nullptr;
    long s_entries = 0;       // padded entries
    long s_idx_bytes = 0;
    bool f_vals64 = false;    // FP32 vectors read the double values
};
// y = A x (T = double: A.val, T = float: A.valf); if eta != 0: y -= eta x;
// if u: *udot_out = u.y (double)
template <typename T>
void spmv(hipStream_t s, const RedBuf &rb, const CsrDev &A, const T *x, T *y, double eta,
          const T *u, double *udot_out);

// Y = A X for k vectors (column-major, leading dimensions ldx, ldy), FP64,
// in passes of up to 8 vectors over the matrix (kernels_spmm.hip)
void spmm(hipStream_t s, const CsrDev &A, int k, const double *x, long ldx, double *y,
          long ldy);

// ---- preconditioners / conversion ----
template <typename T>
void diag_apply(hipStream_t s, T *z, const T *d, const T *r, long n);
void convert(hipStream_t s, float *dst, const double *src, long n);
void convert(hipStream_t s, double *dst, const float *src, long n);

// ---- inner solver (LSOLREVCOM) fused steps, T = double or float ----
// z = P^-1 r is already in z: ycoef = Q^T z ; z -= Q ycoef ; rho = r.z
template <typename T>
void cg_project_rho(hipStream_t s, const RedBuf &rb, T *z, const T *Q, int jq, const T *r,
                    long n, InnerScalars *sc, double *ycoef);
// z += (rho/rho1) p ; p = z   (iteration number from the host)
template <typename T>
void cg_direction(hipStream_t s, T *z, T *p, long n, InnerScalars *sc, int iter);
// q already = (A - eta) p, alprho = u.q ; q -= u alprho ; pq = p.q ; alpha
template <typename T>
void cg_project_alpha(hipStream_t s, const RedBuf &rb, T *q, const T *u, const T *p, long n,
                      InnerScalars *sc);
// PCG: r -= alpha q ; gk2 = r.r ; x += alpha p ; xn2 = x.x
template <typename T>
void cg_update_posdef(hipStream_t s, const RedBuf &rb, T *r, const T *q, T *x, const T *p,
                      long n, InnerScalars *sc);
// sQMR: r -= alpha q ; gk2 ; smoothing scalars ; then x, x2 update
template <typename T>
void qmr_update_r(hipStream_t s, const RedBuf &rb, T *r, const T *q, long n, InnerScalars *sc);
template <typename T>
void qmr_update_x(hipStream_t s, const RedBuf &rb, T *x, T *x2, const T *p, long n,
                  InnerScalars *sc, int iter);

}  // namespace jdhip
