// kernels_spmm.hip -- symmetric SpMM: Y = A X for a block of k vectors.
//
// MATVECA (src/jd/JDaux.F:3397-3418) applied to k vectors at once.  The
// matrix is the device operator of a context (full symmetric pattern built
// from the caller's upper triangle), in either of its two formats:
//  * sliced ELL: one wave64 per slice of 64 rows, one lane per row, the
//    slice's entries column-major so each load of the wave is contiguous;
//  * CSR: 8 lanes per row.
// Every matrix entry (value and column) is loaded once per pass and used for
// KB vectors, so the matrix bytes are amortised over the block: algorithmic
// bytes (s+4) nnz_u + 4 (n+1) + 2 s n k for k vectors, against k times the
// matrix for k separate SpMVs.  Vectors are column-major, X(i, j) =
// x[i + j*ldx]; blocks wider than 8 run in passes of 8.
#include "kernels.hpp"
#include "xcd.hpp"

namespace jdhip {

namespace {

template <int KB>
__global__ __launch_bounds__(kBlock) void k_spmm_sell(int n, int nslices,
                                                      const long *__restrict__ vptr,
                                                      const long *__restrict__ iptr,
                                                      const int *__restrict__ wm,
                                                      const double *__restrict__ val,
                                                      const char *__restrict__ idx,
                                                      const double *__restrict__ x, long ldx,
                                                      double *__restrict__ y, long ldy) {
    constexpr int WPB = kBlock / 64;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (long sl = logical_block(1) * WPB + wv; sl < nslices; sl += (long)gridDim.x * WPB) {
        const long row = sl * 64 + lane;
        if (row >= n) continue;
        const int code = wm[sl];
        const int w = code >> 1;
        const double *v = val + vptr[sl] + lane;
        double s[KB];
#pragma unroll
        for (int j = 0; j < KB; ++j) s[j] = 0.0;
        if (code & 1) {
            const short *ci = (const short *)(idx + iptr[sl]) + lane;
            for (int k = 0; k < w; ++k) {
                const double a = v[k * 64];
                const long c = row + ci[k * 64];
#pragma unroll
                for (int j = 0; j < KB; ++j) s[j] += a * x[c + j * ldx];
            }
        } else {
            const int *ci = (const int *)(idx + iptr[sl]) + lane;
            for (int k = 0; k < w; ++k) {
                const double a = v[k * 64];
                const long c = ci[k * 64];
#pragma unroll
                for (int j = 0; j < KB; ++j) s[j] += a * x[c + j * ldx];
            }
        }
#pragma unroll
        for (int j = 0; j < KB; ++j) y[row + j * ldy] = s[j];
    }
}

template <int KB>
__global__ __launch_bounds__(kBlock) void k_spmm_csr(int n, const int *__restrict__ rowptr,
                                                     const int *__restrict__ col,
                                                     const double *__restrict__ val,
                                                     const double *__restrict__ x, long ldx,
                                                     double *__restrict__ y, long ldy) {
    constexpr int LPR = 8, GPB = kBlock / LPR;
    const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    for (long row = logical_block(1) * GPB + g; row < n; row += (long)gridDim.x * GPB) {
        double s[KB];
#pragma unroll
        for (int j = 0; j < KB; ++j) s[j] = 0.0;
        for (int k = rowptr[row] + l; k < rowptr[row + 1]; k += LPR) {
            const double a = val[k];
            const long c = col[k];
#pragma unroll
            for (int j = 0; j < KB; ++j) s[j] += a * x[c + j * ldx];
        }
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            double t = s[j];
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, LPR);
            if (l == 0) y[row + j * ldy] = t;
        }
    }
}

template <int KB>
void spmm_pass(hipStream_t s, const CsrDev &A, const double *x, long ldx, double *y, long ldy) {
    if (A.lpr == 0 && A.nslices > 0) {
        const long blocks = ((long)A.nslices + kBlock / 64 - 1) / (kBlock / 64);
        const int g = (int)std::min<long>(blocks, 16384);
        k_spmm_sell<KB><<<g, kBlock, 0, s>>>(A.n, A.nslices, A.s_vptr, A.s_iptr, A.s_wm,
                                             A.s_val, A.s_idx, x, ldx, y, ldy);
    } else {
        const long blocks = ((long)A.n + kBlock / 8 - 1) / (kBlock / 8);
        const int g = (int)std::min<long>(blocks, 16384);
        k_spmm_csr<KB><<<g, kBlock, 0, s>>>(A.n, A.rowptr, A.col, A.val, x, ldx, y, ldy);
    }
}

}  // namespace

void spmm(hipStream_t s, const CsrDev &A, int k, const double *x, long ldx, double *y,
          long ldy) {
    for (int j0 = 0; j0 < k; j0 += 8) {
        const int kb = std::min(8, k - j0);
        const double *xj = x + (long)j0 * ldx;
        double *yj = y + (long)j0 * ldy;
        switch (kb) {
            case 1: spmm_pass<1>(s, A, xj, ldx, yj, ldy); break;
            case 2: spmm_pass<2>(s, A, xj, ldx, yj, ldy); break;
            case 3: spmm_pass<3>(s, A, xj, ldx, yj, ldy); break;
            case 4: spmm_pass<4>(s, A, xj, ldx, yj, ldy); break;
            case 5: spmm_pass<5>(s, A, xj, ldx, yj, ldy); break;
            case 6: spmm_pass<6>(s, A, xj, ldx, yj, ldy); break;
            case 7: spmm_pass<7>(s, A, xj, ldx, yj, ldy); break;
            default: spmm_pass<8>(s, A, xj, ldx, yj, ldy); break;
        }
    }
}

}  // namespace jdhip
