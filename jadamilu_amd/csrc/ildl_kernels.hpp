// ildl_kernels.hpp -- gfx950 kernels of the multilevel ILDL apply.
//
// Per level (src/ilupack/symAMGsol.c:173-628 restated), with the coupling
// block F^T = A_CB (the reference's structure):
//   gather+scale -> y_B = B^-1 b_B (forward L_B sweep, D^-1, backward L_B^T
//   sweep) -> coarse rhs b_C - F^T y_B -> [next level] -> x_B = y_B -
//   B^-1 F x_C (a second solve with B) -> scatter+scale;
// with the factor's own coupling rows L_CB instead:
//   gather+scale -> forward L_B sweep -> coarse rhs (C rows) -> D^-1 ->
//   [next level] -> B rows minus L_CB^T z_C -> backward L_B^T sweep ->
//   scatter+scale.
// The triangular sweeps are level-scheduled (one launch per dependency
// level, rows of a level in parallel) or, in sweep mode, a fixed number of
// Jacobi iterations x <- b - N x on the strictly triangular part.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace jdhip {

// Factor values are T (double or float), work vectors W (double or float);
// sums are double.
template <typename TI, typename W>
void ildl_gather(hipStream_t s, int n, const int *perm, const double *scal, const TI *in,
                 W *buf);
template <typename W, typename TO>
void ildl_scatter(hipStream_t s, int n, int nB, const int *perm, const double *scal,
                  const W *buf, const W *zc, TO *out);
// x[rows of one level] -= sum_j M(row, j) x[j]   (unit-diagonal triangular)
template <typename T, typename W>
void ildl_trsv_level(hipStream_t s, const int *rows, int count, const int *ptr, const int *idx,
                     const T *val, W *x);
// Device view of a SELL-C-sigma matrix: slices of 64 rows (one wave64, one
// lane per row) taken from windows of rows sorted by length; entries of a
// slice stored column-major (value at val[vptr[s] + k*64 + lane], k <
// width); the slice's column indices start at byte iptr[s] of idx, as
// 16-bit offsets from cbase[s] when code[s] is odd, else as int32 columns;
// width = code[s] >> 1; rowid[s*64 + lane] is the row a lane owns (-1 for
// padding lanes).
struct SellView {
    int nslices = 0;
    const long *vptr = nullptr;
    const long *iptr = nullptr;
    const int *code = nullptr;
    const int *rowid = nullptr;
    const int *cbase = nullptr;
    const char *idx = nullptr;
};
// One Jacobi sweep of the unit triangle, xout = b - M xin.  xcd bit 0:
// XCD-contiguous block order; bit 1: nontemporal loads of the matrix stream.
template <typename T, typename W>
void ildl_jacobi_sweep_sell(hipStream_t s, const SellView &M, const T *val, const W *b,
                            const W *xin, W *xout, int xcd);
// CSR forms with lpr (4, 8, 16, 32 or 64) lanes per row, for matrices with long rows;
// xcd != 0: XCD-contiguous block order (as for the SELL and stream forms)
template <typename T, typename W>
void ildl_jacobi_csr(hipStream_t s, int lpr, int nrows, const int *ptr, const int *idx,
                     const T *val, const W *b, const W *xin, W *xout, int xcd);
template <typename T, typename W>
void ildl_sub_csr(hipStream_t s, int lpr, int nrows, const int *ptr, const int *idx,
                  const T *val, const W *x, W *y, int xcd);
// y[r] -= sum_j M(r, j) x[j] over the rows of a SELL-C-sigma matrix
template <typename T, typename W>
void ildl_sub_sell(hipStream_t s, const SellView &M, const T *val, const W *x, W *y, int xcd);
// CSR-stream form: blocks of consecutive rows with at most 2048 entries
// (and 256 rows) each, or one longer row; rblk[b] = first row of block b,
// rblk[nblk] = nrows.  Entries are streamed through LDS with coalesced loads.
// xcd != 0: every XCD takes one contiguous range of row blocks.
std::vector<int> stream_row_blocks(int nrows, const int *ptr);
template <typename T, typename W>
void ildl_jacobi_stream(hipStream_t s, int nblk, const int *rblk, const int *ptr,
                        const int *idx, const T *val, const W *b, const W *xin, W *xout,
                        int xcd);
template <typename T, typename W>
void ildl_sub_stream(hipStream_t s, int nblk, const int *rblk, const int *ptr, const int *idx,
                     const T *val, const W *x, W *y, int xcd);
template <typename W>
void ildl_dscale(hipStream_t s, int n, const double *dinv, W *x);
// x = D^-1 x with 2x2 blocks: row i pairs with pair[i] (-1: 1x1 pivot,
// x_i *= dinv_i); a pair is updated by its lower index,
// (x_i, x_p) <- (dinv_i x_i + doff_i x_p, doff_p x_i + dinv_p x_p)
template <typename W>
void ildl_dscale2(hipStream_t s, int n, const double *dinv, const double *doff, const int *pair,
                  W *x);
// x = 0 (a kernel, not a memset: the apply may be captured in a HIP graph)
template <typename W>
void ildl_zero(hipStream_t s, int n, W *x);
// y += t
template <typename W>
void ildl_add(hipStream_t s, int n, const W *t, W *y);
// y = Minv x, Minv m x m row-major
template <typename T, typename W>
void ildl_dense_gemv(hipStream_t s, int m, const T *Minv, const W *x, W *y);

}  // namespace jdhip
