// kernels_inner.hip -- inner correction-equation solver (LSOLREVCOM,
// src/jd/JDaux.F:2989-3268) as fused gfx950 kernels, templated on the vector
// type T: double for the FP64 path, float for the mixed path (FP32 vectors,
// FP32 SpMV and preconditioner; every dot product and norm accumulates in
// double).  All scalars live in device memory (InnerScalars), so a step is a
// chain of launches with no host round trip; breakdown tests set a flag
// that turns the rest of the step into no-ops (the reference's early exits).
#include "kernels.hpp"
#include "reduce.hpp"
#include "rowsum.hpp"
#include "xcd.hpp"

#include <algorithm>

namespace jdhip {

// ---------------------------------------------------------------------------
// out[c] = sum_i V[c*ldv+i] * w[i]
template <typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_multidot(const T *__restrict__ V, long ldv, int jv,
                                                     const T *__restrict__ w, long n,
                                                     double *partials, int nparts,
                                                     double *out) {
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        const double wi = (double)w[i];
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (c < jv) acc[c] += (double)V[c * ldv + i] * wi;
    }
    __shared__ double fin[NC];
    if (grid_reduce<NC>(acc, jv, partials, nparts, fin)) {
        if ((int)threadIdx.x < jv) out[threadIdx.x] = fin[threadIdx.x];
    }
}

// finish of a multi-column reduction: block c sums column c of the partials
// (columns are independent, so they are summed in parallel)
__global__ __launch_bounds__(kBlock) void k_finish_columns(const double *__restrict__ partials,
                                                          int nparts, double *out) {
    __shared__ double red[kBlock / 64];
    const int c = blockIdx.x, lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const double s = wave_sum(thread_partials(partials + (long)c * nparts, nparts));
    if (lane == 0) red[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) t += red[w];
        out[c] = t;
    }
}

template <typename T>
void multidot(hipStream_t s, const RedBuf &rb, const T *V, long ldv, int jv, const T *w, long n,
              double *out) {
    if (jv <= 0) return;
    const int g = grid_for(n);
    double *P = rb.partials;
#define JDHIP_MD(NC)                                                                    \
    do {                                                                                \
        k_multidot<T, NC><<<g, kBlock, 0, s>>>(V, ldv, jv, w, n, P, 0, out);            \
        k_finish_columns<<<jv, kBlock, 0, s>>>(P, g, out);                               \
    } while (0)
    if (jv <= 1) JDHIP_MD(1);
    else if (jv <= 2) JDHIP_MD(2);
    else if (jv <= 4) JDHIP_MD(4);
    else if (jv <= 8) JDHIP_MD(8);
    else if (jv <= 16) JDHIP_MD(16);
    else if (jv <= 32) JDHIP_MD(32);
    else JDHIP_MD(64);
#undef JDHIP_MD
}

// ---------------------------------------------------------------------------
// SpMV on the device-built full CSR; LPR lanes per row; optional epilogue
//   y_i -= eta x_i   and   udot += u_i y_i   (u.y in double).
// Rows are accumulated in double also for FP32 storage: near an eigenvalue
// of A - eta I the row sums cancel to a tiny fraction of their terms, and a
// float accumulator would leave errors of the size of that result.
template <typename T, typename TA, int LPR, bool ETA, bool UDOT>
__global__ __launch_bounds__(kBlock) void k_spmv(int n, const int *__restrict__ rowptr,
                                                 const int *__restrict__ col,
                                                 const TA *__restrict__ val,
                                                 const T *__restrict__ x, T *__restrict__ y,
                                                 double eta_host, const T *__restrict__ u,
                                                 double *partials, int nparts,
                                                 double *udot_out) {
    constexpr int GPB = kBlock / LPR;
    const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    double acc[1] = {0.0};
    for (long row = logical_block(1) * GPB + g; row < n; row += (long)gridDim.x * GPB) {
        const int b = rowptr[row], e = rowptr[row + 1];
        double s = csr_lane_sum<LPR>(col, val, x, b, e, l);
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
        if (l == 0) {
            if (ETA) s -= eta_host * (double)x[row];
            const T yi = (T)s;
            y[row] = yi;
            if (UDOT) acc[0] += (double)u[row] * (double)yi;
        }
    }
    if (UDOT) {
        __shared__ double fin[1];
        if (grid_reduce<1>(acc, 1, partials, nparts, fin))
            if (threadIdx.x == 0) udot_out[0] = fin[0];
    }
}

template <typename T, typename TA, int LPR>
static void spmv_lpr(hipStream_t s, const RedBuf &rb, int n, const int *rp, const int *ci,
                     const TA *val, const T *x, T *y, double eta, const T *u, double *udot) {
    constexpr int GPB = kBlock / LPR;
    long rows_blocks = ((long)n + GPB - 1) / GPB;
    const bool has_eta = eta != 0.0;
    double *P = rb.partials;
    if (u) {
        int g = (int)std::min<long>(rows_blocks, kMaxRedBlocks);
        if (has_eta) {
            k_spmv<T, TA, LPR, true, true><<<g, kBlock, 0, s>>>(n, rp, ci, val, x, y, eta, u, P, 0,
                                                           udot);
            k_spmv<T, TA, LPR, true, true><<<1, kBlock, 0, s>>>(0, rp, ci, val, x, y, eta, u, P, g,
                                                           udot);
        } else {
            k_spmv<T, TA, LPR, false, true><<<g, kBlock, 0, s>>>(n, rp, ci, val, x, y, eta, u, P, 0,
                                                            udot);
            k_spmv<T, TA, LPR, false, true><<<1, kBlock, 0, s>>>(0, rp, ci, val, x, y, eta, u, P, g,
                                                            udot);
        }
    } else {
        int g = (int)std::min<long>(rows_blocks, 8192);
        if (has_eta)
            k_spmv<T, TA, LPR, true, false><<<g, kBlock, 0, s>>>(n, rp, ci, val, x, y, eta, u, P, 0,
                                                            udot);
        else
            k_spmv<T, TA, LPR, false, false><<<g, kBlock, 0, s>>>(n, rp, ci, val, x, y, eta, u, P, 0,
                                                             udot);
    }
}

// Sliced ELL: one wave64 per slice of 64 rows, lane = row within the slice.
template <typename T, typename TA, bool ETA, bool UDOT>
__global__ __launch_bounds__(kBlock) void k_spmv_sell(
    int n, int nslices, const long *__restrict__ vptr, const long *__restrict__ iptr,
    const int *__restrict__ wm, const TA *__restrict__ val, const char *__restrict__ idx,
    const T *__restrict__ x, T *__restrict__ y, double eta_host, const T *__restrict__ u,
    double *partials, int nparts, double *udot_out) {
    constexpr int WPB = kBlock / 64;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double acc[1] = {0.0};
    // XCD-contiguous slice ranges: a slice's x gathers overlap its
    // neighbours', so they are served from one L2
    for (long sl = logical_block(1) * WPB + wv; sl < nslices; sl += (long)gridDim.x * WPB) {
        const long row = sl * 64 + lane;
        if (row >= n) continue;
        const int code = wm[sl];
        const int w = code >> 1;
        const TA *v = val + vptr[sl] + lane;
        double s0 = 0.0, s1 = 0.0;
        if (code & 1) {
            const short *ci = (const short *)(idx + iptr[sl]) + lane;
            int k = 0;
            for (; k + 1 < w; k += 2) {
                s0 += (double)v[k * 64] * (double)x[row + ci[k * 64]];
                s1 += (double)v[(k + 1) * 64] * (double)x[row + ci[(k + 1) * 64]];
            }
            if (k < w) s0 += (double)v[k * 64] * (double)x[row + ci[k * 64]];
        } else {
            const int *ci = (const int *)(idx + iptr[sl]) + lane;
            int k = 0;
            for (; k + 1 < w; k += 2) {
                s0 += (double)v[k * 64] * (double)x[ci[k * 64]];
                s1 += (double)v[(k + 1) * 64] * (double)x[ci[(k + 1) * 64]];
            }
            if (k < w) s0 += (double)v[k * 64] * (double)x[ci[k * 64]];
        }
        double ys = s0 + s1;
        if (ETA) ys -= eta_host * (double)x[row];
        const T yi = (T)ys;
        y[row] = yi;
        if (UDOT) acc[0] += (double)u[row] * (double)yi;
    }
    if (UDOT) {
        __shared__ double fin[1];
        if (grid_reduce<1>(acc, 1, partials, nparts, fin))
            if (threadIdx.x == 0) udot_out[0] = fin[0];
    }
}

template <typename T, typename TA>
static void spmv_sell(hipStream_t s, const RedBuf &rb, const CsrDev &A, const TA *val,
                      const T *x, T *y, double eta, const T *u, double *udot) {
    constexpr int WPB = kBlock / 64;
    const long blocks = ((long)A.nslices + WPB - 1) / WPB;
    const bool has_eta = eta != 0.0;
    double *P = rb.partials;
    if (u) {
        // one partial sum per block: the full grid fits the scratch
        int g = (int)std::min<long>(blocks, kMaxPartials);
        if (has_eta) {
            k_spmv_sell<T, TA, true, true><<<g, kBlock, 0, s>>>(A.n, A.nslices, A.s_vptr, A.s_iptr,
                                                          A.s_wm, val, A.s_idx, x, y, eta, u,
                                                          P, 0, udot);
            k_spmv_sell<T, TA, true, true><<<1, kBlock, 0, s>>>(A.n, 0, A.s_vptr, A.s_iptr,
                                                          A.s_wm, val, A.s_idx, x, y, eta, u,
                                                          P, g, udot);
        } else {
            k_spmv_sell<T, TA, false, true><<<g, kBlock, 0, s>>>(A.n, A.nslices, A.s_vptr,
                                                           A.s_iptr, A.s_wm, val, A.s_idx, x, y,
                                                           eta, u, P, 0, udot);
            k_spmv_sell<T, TA, false, true><<<1, kBlock, 0, s>>>(A.n, 0, A.s_vptr, A.s_iptr,
                                                           A.s_wm, val, A.s_idx, x, y, eta, u,
                                                           P, g, udot);
        }
    } else {
        int g = (int)std::min<long>(blocks, 1L << 20);
        if (has_eta)
            k_spmv_sell<T, TA, true, false><<<g, kBlock, 0, s>>>(A.n, A.nslices, A.s_vptr,
                                                           A.s_iptr, A.s_wm, val, A.s_idx, x, y,
                                                           eta, u, P, 0, udot);
        else
            k_spmv_sell<T, TA, false, false><<<g, kBlock, 0, s>>>(A.n, A.nslices, A.s_vptr,
                                                            A.s_iptr, A.s_wm, val, A.s_idx, x,
                                                            y, eta, u, P, 0, udot);
    }
}

template <typename T, typename TA>
static void spmv_vals(hipStream_t s, const RedBuf &rb, const CsrDev &A, const TA *val,
                      const TA *sv, const T *x, T *y, double eta, const T *u,
                      double *udot_out) {
    if (A.lpr == 0 && A.nslices > 0) {
        spmv_sell<T, TA>(s, rb, A, sv, x, y, eta, u, udot_out);
        return;
    }
    switch (A.lpr) {
    case 1: spmv_lpr<T, TA, 1>(s, rb, A.n, A.rowptr, A.col, val, x, y, eta, u, udot_out); break;
    case 2: spmv_lpr<T, TA, 2>(s, rb, A.n, A.rowptr, A.col, val, x, y, eta, u, udot_out); break;
    case 4: spmv_lpr<T, TA, 4>(s, rb, A.n, A.rowptr, A.col, val, x, y, eta, u, udot_out); break;
    case 16: spmv_lpr<T, TA, 16>(s, rb, A.n, A.rowptr, A.col, val, x, y, eta, u, udot_out); break;
    case 32: spmv_lpr<T, TA, 32>(s, rb, A.n, A.rowptr, A.col, val, x, y, eta, u, udot_out); break;
    default: spmv_lpr<T, TA, 8>(s, rb, A.n, A.rowptr, A.col, val, x, y, eta, u, udot_out); break;
    }
}

template <typename T>
void spmv(hipStream_t s, const RedBuf &rb, const CsrDev &A, const T *x, T *y, double eta,
          const T *u, double *udot_out) {
    if constexpr (sizeof(T) == 8)
        spmv_vals<T, double>(s, rb, A, A.val, A.s_val, x, y, eta, u, udot_out);
    else if (A.f_vals64)
        spmv_vals<T, double>(s, rb, A, A.val, A.s_val, x, y, eta, u, udot_out);
    else
        spmv_vals<T, float>(s, rb, A, A.valf, A.s_valf, x, y, eta, u, udot_out);
}

// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag(T *z, const T *d, const T *r, long n) {
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock)
        z[i] = d[i] * r[i];
}
template <typename T>
void diag_apply(hipStream_t s, T *z, const T *d, const T *r, long n) {
    k_diag<T><<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(z, d, r, n);
}

template <typename TD, typename TS>
__global__ __launch_bounds__(kBlock) void k_convert(TD *dst, const TS *src, long n) {
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock)
        dst[i] = (TD)src[i];
}
void convert(hipStream_t s, float *dst, const double *src, long n) {
    if (n > 0) k_convert<float, double><<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(dst, src, n);
}
void convert(hipStream_t s, double *dst, const float *src, long n) {
    if (n > 0) k_convert<double, float><<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(dst, src, n);
}

// ---------------------------------------------------------------------------
// After the preconditioner: z -= Q (Q^T z) ; rho = r.z ; breakdown test
// (JDaux.F:1916-1922 projection, then LSOLREVCOM label 3).
template <typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_cg_project_rho(T *__restrict__ z,
                                                           const T *__restrict__ Q, int jq,
                                                           const T *__restrict__ r, long n,
                                                           InnerScalars *sc,
                                                           const double *__restrict__ ycoef,
                                                           double *partials, int nparts) {
    if (sc->flag) return;
    // the projection cancels the large components of z along the basis:
    // summed in double also for FP32 vectors
    __shared__ double ys[NC];
    if ((int)threadIdx.x < jq) ys[threadIdx.x] = -ycoef[threadIdx.x];
    __syncthreads();
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        double zd = (double)z[i];
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (c < jq) zd += ys[c] * (double)Q[c * n + i];
        const T zi = (T)zd;
        z[i] = zi;
        acc[0] += (double)r[i] * (double)zi;
    }
    __shared__ double fin[1];
    if (grid_reduce<1>(acc, 1, partials, nparts, fin)) {
        if (threadIdx.x == 0) {
            double rho = fin[0];
            sc->rho = rho;
            if ((sc->posdef && rho <= 0.0) || rho == 0.0) sc->flag = 1;
        }
    }
}

template <typename T>
void cg_project_rho(hipStream_t s, const RedBuf &rb, T *z, const T *Q, int jq, const T *r,
                    long n, InnerScalars *sc, double *ycoef) {
    multidot<T>(s, rb, Q, n, jq, z, n, ycoef);
    const int g = grid_for(n);
    double *p = rb.partials;
#define JDHIP_PR(NC)                                                                    \
    do {                                                                                \
        k_cg_project_rho<T, NC><<<g, kBlock, 0, s>>>(z, Q, jq, r, n, sc, ycoef, p, 0);  \
        k_cg_project_rho<T, NC><<<1, kBlock, 0, s>>>(z, Q, jq, r, 0, sc, ycoef, p, g);  \
    } while (0)
    if (jq <= 1) JDHIP_PR(1);
    else if (jq <= 2) JDHIP_PR(2);
    else if (jq <= 4) JDHIP_PR(4);
    else if (jq <= 8) JDHIP_PR(8);
    else if (jq <= 16) JDHIP_PR(16);
    else if (jq <= 32) JDHIP_PR(32);
    else JDHIP_PR(64);
#undef JDHIP_PR
}

// direction (LSOLREVCOM label 3): z += (rho/rho1) p ; p = z
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_direction(T *__restrict__ z, T *__restrict__ p,
                                                         long n, const InnerScalars *sc,
                                                         int iter) {
    if (sc->flag) return;
    const bool first = iter <= 1;
    const T beta = first ? T(0) : (T)(sc->rho / sc->rho1);
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        T zi = z[i];
        if (!first && beta != T(0)) zi += beta * p[i];
        z[i] = zi;
        p[i] = zi;
    }
}
template <typename T>
void cg_direction(hipStream_t s, T *z, T *p, long n, InnerScalars *sc, int iter) {
    k_cg_direction<T><<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(z, p, n, sc, iter);
}

// after the matvec (JDaux.F:1898-1909 and LSOLREVCOM label 4):
// q -= u alprho ; pq = p.q ; alpha = rho / pq
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_project_alpha(T *__restrict__ q,
                                                             const T *__restrict__ u,
                                                             const T *__restrict__ p, long n,
                                                             InnerScalars *sc, double *partials,
                                                             int nparts) {
    if (sc->flag) return;
    const T a = (T)(-sc->alprho);
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        T qi = q[i] + a * u[i];
        q[i] = qi;
        acc[0] += (double)p[i] * (double)qi;
    }
    __shared__ double fin[1];
    if (grid_reduce<1>(acc, 1, partials, nparts, fin)) {
        if (threadIdx.x == 0) {
            double pq = fin[0];
            sc->pq = pq;
            if ((sc->posdef && pq <= 0.0) || pq == 0.0) sc->flag = 2;
            else sc->alpha = sc->rho / pq;
        }
    }
}
template <typename T>
void cg_project_alpha(hipStream_t s, const RedBuf &rb, T *q, const T *u, const T *p, long n,
                      InnerScalars *sc) {
    const int g = grid_for(n);
    k_cg_project_alpha<T><<<g, kBlock, 0, s>>>(q, u, p, n, sc, rb.partials, 0);
    k_cg_project_alpha<T><<<1, kBlock, 0, s>>>(q, u, p, 0, sc, rb.partials, g);
}

// PCG update: r -= alpha q ; x += alpha p ; ||r||^2, ||x||^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update_posdef(T *__restrict__ r,
                                                             const T *__restrict__ q,
                                                             T *__restrict__ x,
                                                             const T *__restrict__ p, long n,
                                                             InnerScalars *sc, double *partials,
                                                             int nparts) {
    if (sc->flag) return;
    const T al = (T)sc->alpha;
    double acc[2] = {0.0, 0.0};
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        T ri = r[i] + (-al) * q[i];
        r[i] = ri;
        acc[0] += (double)ri * (double)ri;
        T xi = x[i] + al * p[i];
        x[i] = xi;
        acc[1] += (double)xi * (double)xi;
    }
    __shared__ double fin[2];
    if (grid_reduce<2>(acc, 2, partials, nparts, fin)) {
        if (threadIdx.x == 0) {
            sc->gk2 = fin[0];
            sc->xn2 = fin[1];
            sc->rho1 = sc->rho;
        }
    }
}
template <typename T>
void cg_update_posdef(hipStream_t s, const RedBuf &rb, T *r, const T *q, T *x, const T *p,
                      long n, InnerScalars *sc) {
    const int g = grid_for(n);
    k_cg_update_posdef<T><<<g, kBlock, 0, s>>>(r, q, x, p, n, sc, rb.partials, 0);
    k_cg_update_posdef<T><<<1, kBlock, 0, s>>>(r, q, x, p, 0, sc, rb.partials, g);
}

// sQMR (LSOLREVCOM, JDaux.F:3167-3190): r -= alpha q, smoothing scalars
template <typename T>
__global__ __launch_bounds__(kBlock) void k_qmr_update_r(T *__restrict__ r,
                                                         const T *__restrict__ q, long n,
                                                         InnerScalars *sc, double *partials,
                                                         int nparts) {
    if (sc->flag) return;
    const T al = (T)sc->alpha;
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        T ri = r[i] + (-al) * q[i];
        r[i] = ri;
        acc[0] += (double)ri * (double)ri;
    }
    __shared__ double fin[1];
    if (grid_reduce<1>(acc, 1, partials, nparts, fin)) {
        if (threadIdx.x == 0) {
            double gk = sqrt(fin[0]);
            sc->gk2 = fin[0];
            double gk2 = gk * gk;
            sc->ttk2m = sc->ttk2;
            sc->ttk2 = gk2 / sc->ggk2;
            sc->ck2 = 1.0 / (1.0 + sc->ttk2);
            sc->ggk2 = sc->ggk2 * sc->ttk2 * sc->ck2;
        }
    }
}
template <typename T>
void qmr_update_r(hipStream_t s, const RedBuf &rb, T *r, const T *q, long n, InnerScalars *sc) {
    const int g = grid_for(n);
    k_qmr_update_r<T><<<g, kBlock, 0, s>>>(r, q, n, sc, rb.partials, 0);
    k_qmr_update_r<T><<<1, kBlock, 0, s>>>(r, q, 0, sc, rb.partials, g);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_qmr_update_x(T *__restrict__ x, T *__restrict__ x2,
                                                         const T *__restrict__ p, long n,
                                                         InnerScalars *sc, int iter,
                                                         double *partials, int nparts) {
    if (sc->flag) return;
    const double al = sc->alpha;
    const bool first = iter <= 1;
    const T c1 = first ? T(0) : (T)(al / sc->ttk2m);
    const T c2 = first ? (T)(al * sc->ck2) : (T)(sc->ttk2m * sc->ck2);
    double acc[1] = {0.0};
    for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        T xi;
        if (first) {
            T t = p[i] * c2;
            x2[i] = t;
            xi = t;
        } else {
            T t = (x2[i] + c1 * p[i]) * c2;
            x2[i] = t;
            xi = x[i] + t;
        }
        x[i] = xi;
        acc[0] += (double)xi * (double)xi;
    }
    __shared__ double fin[1];
    if (grid_reduce<1>(acc, 1, partials, nparts, fin))
        if (threadIdx.x == 0) {
            sc->xn2 = fin[0];
            sc->rho1 = sc->rho;
        }
}
template <typename T>
void qmr_update_x(hipStream_t s, const RedBuf &rb, T *x, T *x2, const T *p, long n,
                  InnerScalars *sc, int iter) {
    const int g = grid_for(n);
    k_qmr_update_x<T><<<g, kBlock, 0, s>>>(x, x2, p, n, sc, iter, rb.partials, 0);
    k_qmr_update_x<T><<<1, kBlock, 0, s>>>(x, x2, p, 0, sc, iter, rb.partials, g);
}

#define JDHIP_INST(T)                                                                          \
    template void multidot<T>(hipStream_t, const RedBuf &, const T *, long, int, const T *,    \
                              long, double *);                                                 \
    template void spmv<T>(hipStream_t, const RedBuf &, const CsrDev &, const T *, T *, double, \
                          const T *, double *);                                                \
    template void diag_apply<T>(hipStream_t, T *, const T *, const T *, long);                 \
    template void cg_project_rho<T>(hipStream_t, const RedBuf &, T *, const T *, int,          \
                                    const T *, long, InnerScalars *, double *);                \
    template void cg_direction<T>(hipStream_t, T *, T *, long, InnerScalars *, int);           \
    template void cg_project_alpha<T>(hipStream_t, const RedBuf &, T *, const T *, const T *,  \
                                      long, InnerScalars *);                                   \
    template void cg_update_posdef<T>(hipStream_t, const RedBuf &, T *, const T *, T *,        \
                                      const T *, long, InnerScalars *);                        \
    template void qmr_update_r<T>(hipStream_t, const RedBuf &, T *, const T *, long,           \
                                  InnerScalars *);                                             \
    template void qmr_update_x<T>(hipStream_t, const RedBuf &, T *, T *, const T *, long,      \
                                  InnerScalars *, int);
JDHIP_INST(double)
JDHIP_INST(float)

}  // namespace jdhip
