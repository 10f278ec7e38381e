// ildl_kernels.hip -- gfx950 kernels of the multilevel ILDL apply
// (see ildl_kernels.hpp).  Products with short rows run one lane per row
// on SELL-C-sigma copies (every load of a wave contiguous); long rows run
// on CSR with lane groups of 4 to 64 lanes and a butterfly reduction; the exact
// level-scheduled solve uses 4-lane groups per row of a level.  Every loop
// is grid-stride; vector entries are gathered through the column index.
// The factor's values are stored and streamed as T (double, or float on
// the mixed path), the apply's work vectors as W; every sum is double.  On
// indefinite matrices the mixed path keeps W = double: the substitutions
// cancel large terms, and float intermediates measurably slow the outer
// iteration there, while the vector traffic is small next to the
// factor's and the extra arithmetic is free in these HBM-bound loops.
#include "ildl_kernels.hpp"
#include "rowsum.hpp"
#include "xcd.hpp"

#include <algorithm>
#include <vector>

namespace jdhip {

static constexpr int kB = 256;
static constexpr int kLpr = 4;

// With xcd != 0 the kernels below take their block id from logical_block()
// (xcd.hpp): every XCD works on one contiguous range of rows or slices.

static inline int blocks_for(long n, int per_block, int cap = 8192) {
    long b = (n + per_block - 1) / per_block;
    return (int)std::max<long>(1, std::min<long>(b, cap));
}

template <typename TI, typename W>
__global__ __launch_bounds__(kB) void k_gather(int n, const int *__restrict__ perm,
                                               const double *__restrict__ scal,
                                               const TI *__restrict__ in,
                                               W *__restrict__ buf) {
    for (long i = blockIdx.x * (long)kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) {
        int p = perm[i];
        buf[i] = (W)(scal[p] * (double)in[p]);
    }
}

template <typename W, typename TO>
__global__ __launch_bounds__(kB) void k_scatter(int n, int nB, const int *__restrict__ perm,
                                                const double *__restrict__ scal,
                                                const W *__restrict__ buf,
                                                const W *__restrict__ zc,
                                                TO *__restrict__ out) {
    for (long i = blockIdx.x * (long)kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) {
        int p = perm[i];
        double v = (i < nB) ? (double)buf[i] : (double)zc[i - nB];
        out[p] = (TO)(scal[p] * v);
    }
}

template <typename T, typename W>
__global__ __launch_bounds__(kB) void k_trsv_level(const int *__restrict__ rows, int count,
                                                   const int *__restrict__ ptr,
                                                   const int *__restrict__ idx,
                                                   const T *__restrict__ val, W *x) {
    const int g = threadIdx.x / kLpr, l = threadIdx.x % kLpr;
    for (long t = (long)blockIdx.x * (kB / kLpr) + g; t < count;
         t += (long)gridDim.x * (kB / kLpr)) {
        const int r = rows[t];
        double s = 0.0;
        for (int k = ptr[r] + l; k < ptr[r + 1]; k += kLpr) s += (double)val[k] * (double)x[idx[k]];
#pragma unroll
        for (int o = kLpr / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kLpr);
        if (l == 0) x[r] = (W)((double)x[r] - s);
    }
}

// Loads of the matrix stream; NT = nontemporal, so the single-use values and
// columns do not push the reused x entries out of L2.
template <bool NT, typename X>
__device__ __forceinline__ X ld(const X *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// Dot product of one SELL lane's row with x: w entries, column-major with
// stride 64; columns as int32 (CI = int, base 0) or as 16-bit offsets from
// the slice's base column (CI = unsigned short).  Four independent chains
// keep four gathers of x in flight per lane.
template <bool NT, typename T, typename W, typename CI>
__device__ __forceinline__ double sell_row(const T *v, const CI *c, int base, int w,
                                           const W *x) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int k = 0;
    for (; k + 3 < w; k += 4) {
        s0 += (double)ld<NT>(v + k * 64) * (double)x[base + (int)ld<NT>(c + k * 64)];
        s1 += (double)ld<NT>(v + (k + 1) * 64) * (double)x[base + (int)ld<NT>(c + (k + 1) * 64)];
        s2 += (double)ld<NT>(v + (k + 2) * 64) * (double)x[base + (int)ld<NT>(c + (k + 2) * 64)];
        s3 += (double)ld<NT>(v + (k + 3) * 64) * (double)x[base + (int)ld<NT>(c + (k + 3) * 64)];
    }
    for (; k < w; ++k) s0 += (double)ld<NT>(v + k * 64) * (double)x[base + (int)ld<NT>(c + k * 64)];
    return (s0 + s1) + (s2 + s3);
}

template <bool NT, typename T, typename W>
__device__ __forceinline__ double sell_slice_row(const SellView &M, const T *__restrict__ val,
                                                 long sl, int lane, const W *x) {
    const int code = M.code[sl];
    const T *v = val + M.vptr[sl] + lane;
    if (code & 1)
        return sell_row<NT>(v, (const unsigned short *)(M.idx + M.iptr[sl]) + lane, M.cbase[sl],
                            code >> 1, x);
    return sell_row<NT>(v, (const int *)(M.idx + M.iptr[sl]) + lane, 0, code >> 1, x);
}

template <typename T, typename W, bool NT>
__global__ __launch_bounds__(kB) void k_jacobi_sell(SellView M, const T *__restrict__ val,
                                                    const W *b, const W *xin, W *xout,
                                                    int xcd) {
    // b and xout may be the same array (the last sweep writes in place: each
    // lane reads b only at its own row before writing that row)
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (long sl = logical_block(xcd) * (kB / 64) + wv; sl < M.nslices;
         sl += (long)gridDim.x * (kB / 64)) {
        const int r = M.rowid[sl * 64 + lane];
        if (r < 0) continue;
        const double s = sell_slice_row<NT>(M, val, sl, lane, xin);
        xout[r] = (W)((double)b[r] - s);
    }
}

template <typename T, typename W>
__global__ __launch_bounds__(kB) void k_sub_sell(SellView M, const T *__restrict__ val,
                                                 const W *__restrict__ x, W *__restrict__ y,
                                                 int xcd) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (long sl = logical_block(xcd) * (kB / 64) + wv; sl < M.nslices;
         sl += (long)gridDim.x * (kB / 64)) {
        const int r = M.rowid[sl * 64 + lane];
        if (r < 0) continue;
        const double s = sell_slice_row<false>(M, val, sl, lane, x);
        y[r] = (W)((double)y[r] - s);
    }
}

// CSR forms for matrices with long rows (LPR lanes per row, xor-butterfly
// reduction inside the lane group): with one lane per row such levels keep
// too few waves in flight and each lane walks its row serially.
template <typename T, typename W, int LPR>
__global__ __launch_bounds__(kB) void k_jacobi_csr(int nrows, const int *__restrict__ ptr,
                                                   const int *__restrict__ idx,
                                                   const T *__restrict__ val, const W *b,
                                                   const W *xin, W *xout, int xcd) {
    const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    for (long r = logical_block(xcd) * (kB / LPR) + g; r < nrows; r += (long)gridDim.x * (kB / LPR)) {
        double s = csr_lane_sum<LPR>(idx, val, xin, ptr[r], ptr[r + 1], l);
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
        if (l == 0) xout[r] = (W)((double)b[r] - s);
    }
}

template <typename T, typename W, int LPR>
__global__ __launch_bounds__(kB) void k_sub_csr(int nrows, const int *__restrict__ ptr,
                                                const int *__restrict__ idx,
                                                const T *__restrict__ val,
                                                const W *__restrict__ x, W *__restrict__ y,
                                                int xcd) {
    const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    for (long r = logical_block(xcd) * (kB / LPR) + g; r < nrows; r += (long)gridDim.x * (kB / LPR)) {
        double s = csr_lane_sum<LPR>(idx, val, x, ptr[r], ptr[r + 1], l);
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
        if (l == 0) y[r] = (W)((double)y[r] - s);
    }
}

// CSR-stream form: block b owns the rows rblk[b] .. rblk[b+1]-1, chosen on
// the host so that their entries number at most kStreamCap (and rows at most
// kB), or a single longer row.  The block streams the entries of its rows
// (consecutive in CSR) with fully coalesced, independent loads, parks the
// products in LDS, then sums each row with a group of t lanes (t = the power
// of two that spreads the block's rows over its 256 threads, at most 64) and
// a butterfly.  A single row longer than kStreamCap is summed by the whole
// block with strided loads.  Every row of the range is written.
static constexpr int kStreamCap = 2048;

template <typename T, typename W>
__device__ __forceinline__ void stream_rows(long blk, const int *__restrict__ rblk,
                                            const int *__restrict__ ptr,
                                            const int *__restrict__ idx,
                                            const T *__restrict__ val,
                                            const W *__restrict__ x, int &row, double &sum,
                                            bool &owner) {
    __shared__ double prod[kStreamCap];
    __shared__ double red[kB / 64];
    const int tid = threadIdx.x;
    const int r0 = rblk[blk], r1 = rblk[blk + 1];
    const int e0 = ptr[r0], e1 = ptr[r1];
    const int ne = e1 - e0;
    owner = false;
    if (r1 - r0 == 1 && ne > kStreamCap) {  // one long row
        double s0 = 0.0, s1 = 0.0;
        int k = e0 + tid;
        for (; k + kB < e1; k += 2 * kB) {
            s0 += (double)val[k] * (double)x[idx[k]];
            s1 += (double)val[k + kB] * (double)x[idx[k + kB]];
        }
        if (k < e1) s0 += (double)val[k] * (double)x[idx[k]];
        double s = s0 + s1;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < kB / 64; ++w) t += red[w];
            row = r0;
            sum = t;
            owner = true;
        }
        return;
    }
    // ne <= kStreamCap and r1 - r0 <= kB by construction of rblk.  A thread
    // takes the entries tid, tid + kB, ...: all their column indices are
    // loaded first, then the values and the gathers of x, so that each wave
    // has up to 3 x kPer loads in flight rather than one index and one
    // gather at a time.  Positions past the block's entries load its last
    // entry again (in bounds, branch-free) and are not stored.
    constexpr int kPer = kStreamCap / kB;
    static_assert(kPer * kB == kStreamCap, "stream blocks hold kPer entries per thread");
    if (ne > 0) {
        const int last = ne - 1;
        int ci[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) ci[u] = idx[e0 + min(tid + u * kB, last)];
        double pv[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u)
            pv[u] = (double)val[e0 + min(tid + u * kB, last)] * (double)x[ci[u]];
#pragma unroll
        for (int u = 0; u < kPer; ++u)
            if (tid + u * kB < ne) prod[tid + u * kB] = pv[u];
    }
    __syncthreads();
    const int nr = r1 - r0;
    int t = 64;
    while (t > 1 && t * nr > kB) t >>= 1;
    const int g = tid / t, l = tid % t;
    double s = 0.0;
    if (g < nr) {
        const int a = ptr[r0 + g] - e0, b = ptr[r0 + g + 1] - e0;
        for (int k = a + l; k < b; k += t) s += prod[k];
    }
    for (int o = t / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, t);
    if (g < nr && l == 0) {
        row = r0 + g;
        sum = s;
        owner = true;
    }
}

template <typename T, typename W>
__global__ __launch_bounds__(kB) void k_jacobi_stream(const int *__restrict__ rblk,
                                                      const int *__restrict__ ptr,
                                                      const int *__restrict__ idx,
                                                      const T *__restrict__ val, const W *b,
                                                      const W *xin, W *xout, int xcd) {
    int row = 0;
    double s = 0.0;
    bool owner = false;
    stream_rows(logical_block(xcd), rblk, ptr, idx, val, xin, row, s, owner);
    if (owner) xout[row] = (W)((double)b[row] - s);
}

template <typename T, typename W>
__global__ __launch_bounds__(kB) void k_sub_stream(const int *__restrict__ rblk,
                                                   const int *__restrict__ ptr,
                                                   const int *__restrict__ idx,
                                                   const T *__restrict__ val,
                                                   const W *__restrict__ x, W *__restrict__ y,
                                                   int xcd) {
    int row = 0;
    double s = 0.0;
    bool owner = false;
    stream_rows(logical_block(xcd), rblk, ptr, idx, val, x, row, s, owner);
    if (owner) y[row] = (W)((double)y[row] - s);
}

template <typename W>
__global__ __launch_bounds__(kB) void k_dscale(int n, const double *__restrict__ dinv, W *x) {
    for (long i = blockIdx.x * (long)kB + threadIdx.x; i < n; i += (long)gridDim.x * kB)
        x[i] = (W)((double)x[i] * dinv[i]);
}

template <typename W>
__global__ __launch_bounds__(kB) void k_dscale2(int n, const double *__restrict__ dinv,
                                                const double *__restrict__ doff,
                                                const int *__restrict__ pair, W *x) {
    for (long i = blockIdx.x * (long)kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) {
        const int p = pair[i];
        if (p < 0) {
            x[i] = (W)((double)x[i] * dinv[i]);
        } else if (i < p) {
            const double xi = (double)x[i], xp = (double)x[p];
            x[i] = (W)(dinv[i] * xi + doff[i] * xp);
            x[p] = (W)(doff[p] * xi + dinv[p] * xp);
        }
    }
}

template <typename W>
__global__ __launch_bounds__(kB) void k_zero(int n, W *x) {
    for (long i = blockIdx.x * (long)kB + threadIdx.x; i < n; i += (long)gridDim.x * kB)
        x[i] = W(0);
}
template <typename W>
__global__ __launch_bounds__(kB) void k_add(int n, const W *__restrict__ t, W *__restrict__ y) {
    for (long i = blockIdx.x * (long)kB + threadIdx.x; i < n; i += (long)gridDim.x * kB)
        y[i] = (W)((double)y[i] + (double)t[i]);
}

// one wave per row of the dense inverse
template <typename T, typename W>
__global__ __launch_bounds__(kB) void k_dense_gemv(int m, const T *__restrict__ Minv,
                                                   const W *__restrict__ x, W *__restrict__ y) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (long r = (long)blockIdx.x * (kB / 64) + w; r < m; r += (long)gridDim.x * (kB / 64)) {
        const T *row = Minv + r * (long)m;
        double s = 0.0;
        for (int j = lane; j < m; j += 64) s += (double)row[j] * (double)x[j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) y[r] = (W)s;
    }
}

template <typename TI, typename W>
void ildl_gather(hipStream_t s, int n, const int *perm, const double *scal, const TI *in,
                 W *buf) {
    if (n <= 0) return;
    k_gather<TI, W><<<blocks_for(n, kB), kB, 0, s>>>(n, perm, scal, in, buf);
}
template <typename W, typename TO>
void ildl_scatter(hipStream_t s, int n, int nB, const int *perm, const double *scal,
                  const W *buf, const W *zc, TO *out) {
    if (n <= 0) return;
    k_scatter<W, TO><<<blocks_for(n, kB), kB, 0, s>>>(n, nB, perm, scal, buf, zc, out);
}
template <typename T, typename W>
void ildl_trsv_level(hipStream_t s, const int *rows, int count, const int *ptr, const int *idx,
                     const T *val, W *x) {
    if (count <= 0) return;
    k_trsv_level<T, W><<<blocks_for(count, kB / kLpr), kB, 0, s>>>(rows, count, ptr, idx, val,
                                                                   x);
}
template <typename T, typename W>
void ildl_jacobi_sweep_sell(hipStream_t s, const SellView &M, const T *val, const W *b,
                            const W *xin, W *xout, int xcd) {
    if (M.nslices <= 0) return;
    const int g = blocks_for(M.nslices, kB / 64, 1 << 20);
    if (xcd & 2)
        k_jacobi_sell<T, W, true><<<g, kB, 0, s>>>(M, val, b, xin, xout, xcd & 1);
    else
        k_jacobi_sell<T, W, false><<<g, kB, 0, s>>>(M, val, b, xin, xout, xcd & 1);
}
template <typename T, typename W>
void ildl_sub_sell(hipStream_t s, const SellView &M, const T *val, const W *x, W *y, int xcd) {
    if (M.nslices <= 0) return;
    k_sub_sell<T, W><<<blocks_for(M.nslices, kB / 64, 1 << 20), kB, 0, s>>>(M, val, x, y, xcd);
}
template <typename T, typename W>
void ildl_jacobi_csr(hipStream_t s, int lpr, int nrows, const int *ptr, const int *idx,
                     const T *val, const W *b, const W *xin, W *xout, int xcd) {
    if (nrows <= 0) return;
#define JDHIP_JC(L)                                                                     \
    k_jacobi_csr<T, W, L><<<blocks_for(nrows, kB / L, 1 << 20), kB, 0, s>>>(nrows, ptr, idx, \
                                                                            val, b, xin, xout, xcd)
    if (lpr >= 64) JDHIP_JC(64);
    else if (lpr >= 32) JDHIP_JC(32);
    else if (lpr >= 16) JDHIP_JC(16);
    else if (lpr >= 8) JDHIP_JC(8);
    else JDHIP_JC(4);
#undef JDHIP_JC
}
template <typename T, typename W>
void ildl_sub_csr(hipStream_t s, int lpr, int nrows, const int *ptr, const int *idx,
                  const T *val, const W *x, W *y, int xcd) {
    if (nrows <= 0) return;
#define JDHIP_SC(L)                                                                      \
    k_sub_csr<T, W, L><<<blocks_for(nrows, kB / L, 1 << 20), kB, 0, s>>>(nrows, ptr, idx, val, \
                                                                         x, y, xcd)
    if (lpr >= 64) JDHIP_SC(64);
    else if (lpr >= 32) JDHIP_SC(32);
    else if (lpr >= 16) JDHIP_SC(16);
    else if (lpr >= 8) JDHIP_SC(8);
    else JDHIP_SC(4);
#undef JDHIP_SC
}
template <typename T, typename W>
void ildl_jacobi_stream(hipStream_t s, int nblk, const int *rblk, const int *ptr,
                        const int *idx, const T *val, const W *b, const W *xin, W *xout,
                        int xcd) {
    if (nblk <= 0) return;
    k_jacobi_stream<T, W><<<nblk, kB, 0, s>>>(rblk, ptr, idx, val, b, xin, xout, xcd);
}
template <typename T, typename W>
void ildl_sub_stream(hipStream_t s, int nblk, const int *rblk, const int *ptr, const int *idx,
                     const T *val, const W *x, W *y, int xcd) {
    if (nblk <= 0) return;
    k_sub_stream<T, W><<<nblk, kB, 0, s>>>(rblk, ptr, idx, val, x, y, xcd);
}
std::vector<int> stream_row_blocks(int nrows, const int *ptr) {
    std::vector<int> rb;
    rb.push_back(0);
    int r = 0;
    while (r < nrows) {
        const int len = ptr[r + 1] - ptr[r];
        if (len > kStreamCap) {  // a long row alone
            ++r;
            rb.push_back(r);
            continue;
        }
        int e = 0, cnt = 0;
        while (r < nrows && cnt < kB) {
            const int l = ptr[r + 1] - ptr[r];
            if (l > kStreamCap || e + l > kStreamCap) break;
            e += l;
            ++cnt;
            ++r;
        }
        rb.push_back(r);
    }
    return rb;
}

template <typename W>
void ildl_dscale(hipStream_t s, int n, const double *dinv, W *x) {
    if (n <= 0) return;
    k_dscale<W><<<blocks_for(n, kB), kB, 0, s>>>(n, dinv, x);
}
template <typename W>
void ildl_dscale2(hipStream_t s, int n, const double *dinv, const double *doff, const int *pair,
                  W *x) {
    if (n <= 0) return;
    k_dscale2<W><<<blocks_for(n, kB), kB, 0, s>>>(n, dinv, doff, pair, x);
}
template void ildl_dscale2<double>(hipStream_t, int, const double *, const double *, const int *,
                                   double *);
template void ildl_dscale2<float>(hipStream_t, int, const double *, const double *, const int *,
                                  float *);
template <typename W>
void ildl_zero(hipStream_t s, int n, W *x) {
    if (n <= 0) return;
    k_zero<W><<<blocks_for(n, kB), kB, 0, s>>>(n, x);
}
template <typename W>
void ildl_add(hipStream_t s, int n, const W *t, W *y) {
    if (n <= 0) return;
    k_add<W><<<blocks_for(n, kB), kB, 0, s>>>(n, t, y);
}
template void ildl_zero<double>(hipStream_t, int, double *);
template void ildl_zero<float>(hipStream_t, int, float *);
template void ildl_add<double>(hipStream_t, int, const double *, double *);
template void ildl_add<float>(hipStream_t, int, const float *, float *);
template <typename T, typename W>
void ildl_dense_gemv(hipStream_t s, int m, const T *Minv, const W *x, W *y) {
    if (m <= 0) return;
    k_dense_gemv<T, W><<<blocks_for(m, kB / 64, 4096), kB, 0, s>>>(m, Minv, x, y);
}

#define INST(T, W)                                                                           \
    template void ildl_trsv_level<T, W>(hipStream_t, const int *, int, const int *,          \
                                        const int *, const T *, W *);                        \
    template void ildl_jacobi_sweep_sell<T, W>(hipStream_t, const SellView &, const T *,      \
                                               const W *, const W *, W *, int);              \
    template void ildl_sub_sell<T, W>(hipStream_t, const SellView &, const T *, const W *,    \
                                      W *, int);                                             \
    template void ildl_jacobi_csr<T, W>(hipStream_t, int, int, const int *, const int *,      \
                                        const T *, const W *, const W *, W *, int);          \
    template void ildl_sub_csr<T, W>(hipStream_t, int, int, const int *, const int *,         \
                                     const T *, const W *, W *, int);                        \
    template void ildl_dense_gemv<T, W>(hipStream_t, int, const T *, const W *, W *);         \
    template void ildl_jacobi_stream<T, W>(hipStream_t, int, const int *, const int *,        \
                                           const int *, const T *, const W *, const W *,      \
                                           W *, int);                                         \
    template void ildl_sub_stream<T, W>(hipStream_t, int, const int *, const int *,           \
                                        const int *, const T *, const W *, W *, int);
INST(double, double)
INST(float, double)
INST(float, float)
template void ildl_dscale<double>(hipStream_t, int, const double *, double *);
template void ildl_dscale<float>(hipStream_t, int, const double *, float *);
#define INST_IO(A, B)                                                                        \
    template void ildl_gather<A, B>(hipStream_t, int, const int *, const double *, const A *, \
                                    B *);                                                    \
    template void ildl_scatter<B, A>(hipStream_t, int, int, const int *, const double *,      \
                                     const B *, const B *, A *);
INST_IO(double, double)
INST_IO(double, float)
INST_IO(float, double)
INST_IO(float, float)

}  // namespace jdhip
