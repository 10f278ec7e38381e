// kernels_mfma.hip -- the dense contractions of the JD outer loop on the
// FP64 matrix cores of gfx950 (v_mfma_f64_16x16x4_f64).
//
//  * gram:       C (K x M) = V^T W for tall V (n x K), W (n x M): the Gram
//                products of the Rayleigh-Ritz steps (the SYR2K/GEMV 'C'
//                column updates of JDRVCOM, src/jd/JDaux.F:1651-1660,
//                2013-2031, and the cluster Rayleigh-Ritz, 2165-2372).
//  * tall_gemm:  Y (n x J) = V (n x K) Q (K x J): the basis update of SHRINK
//                (DORMQR applied to V and W, JDaux.F:2418-2462) and of the
//                cluster Rayleigh-Ritz.
//
// Both stream n x (K + M) or n x (K + J) doubles once: they are HBM-bound
// (arithmetic intensity 2KM / 8(K+M) flop/byte, below 2 for K, M <= 35),
// so the matrix cores save issue slots and registers, not time; the
// loads are what is tuned.
//
// Lane layout of v_mfma_f64_16x16x4_f64 on a wave of 64 lanes (A 16 x 4,
// B 4 x 16, C/D 16 x 16):  lane l holds A(l%16, l/16), B(l/16, l%16) and
// D(4*r + l/16, l%16) in accumulator element r = 0..3 (measured on gfx950
// with one-hot operands for every lane pair).
#include "kernels.hpp"

namespace jdhip {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int kGB = 256;        // threads per workgroup (4 waves)
constexpr int kGR = 64;         // rows of V and W staged per step
constexpr int kGPad = kGR + 1;  // LDS column stride (doubles): spreads the columns over banks

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Partial Gram of one workgroup: rows r = blockIdx.x*kGR, step gridDim.x*kGR.
// KT, MT: number of 16-wide tiles of K and M.  partials: K*M entries of
// gridDim.x doubles each (entry-major, C(i,j) of block b at
// (i + j*K) * gridDim.x + b), so that the finish reads each entry's partial
// sums contiguously.  The next step's rows are loaded into registers while
// the matrix cores work on the current step (one LDS buffer, register
// double buffering).
template <int KT, int MT>
__global__ __launch_bounds__(kGB) void k_gram(const double *__restrict__ V, long ldv, int K,
                                              const double *__restrict__ W, long ldw, int M,
                                              long n, double *__restrict__ partials) {
    __shared__ double sv[KT * 16 * kGPad];
    __shared__ double sw[MT * 16 * kGPad];
    constexpr int NT = KT * MT;                 // output tiles
    constexpr int TPW = (NT + 3) / 4;           // tiles per wave
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    v4d acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    // staging: thread (wave w, lane q) holds row q of the step for the
    // columns w, w + 4, w + 8, ...; each column's 64 rows are one wave's
    // contiguous 512-byte load
    double rv[KT * 4], rw[MT * 4];
    auto load = [&](long r0) {
        const long row = r0 + lane;
        const bool in = row < n;
#pragma unroll
        for (int j = 0; j < KT * 4; ++j) {
            const int c = wave + 4 * j;
            rv[j] = (in && c < K) ? V[(long)c * ldv + row] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < MT * 4; ++j) {
            const int c = wave + 4 * j;
            rw[j] = (in && c < M) ? W[(long)c * ldw + row] : 0.0;
        }
    };
    const long stride = (long)gridDim.x * kGR;
    long r0 = (long)blockIdx.x * kGR;
    if (r0 < n) load(r0);
    for (; r0 < n; r0 += stride) {
#pragma unroll
        for (int j = 0; j < KT * 4; ++j) sv[(wave + 4 * j) * kGPad + lane] = rv[j];
#pragma unroll
        for (int j = 0; j < MT * 4; ++j) sw[(wave + 4 * j) * kGPad + lane] = rw[j];
        __syncthreads();
        if (r0 + stride < n) load(r0 + stride);
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            const int tile = wave + 4 * t;
            if (tile < NT) {
                const int ti = tile % KT, tj = tile / KT;
                const double *a = sv + (ti * 16 + li) * kGPad + lk;
                const double *b = sw + (tj * 16 + li) * kGPad + lk;
#pragma unroll
                for (int s = 0; s < kGR; s += 4) acc[t] = mfma(a[s], b[s], acc[t]);
            }
        }
        __syncthreads();
    }
    const long nb = gridDim.x;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = wave + 4 * t;
        if (tile < NT) {
            const int ti = tile % KT, tj = tile / KT;
            const int j = tj * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = ti * 16 + 4 * r + lk;
                if (i < K && j < M) partials[(i + (long)j * K) * nb + blockIdx.x] = acc[t][r];
            }
        }
    }
}

// C(i, j) = sum over the blocks of the partials: one wave per output entry,
// lane l adds the blocks l, l + 64, ... in order, then a fixed butterfly over
// the 64 lanes (the same summation order on every call).
__global__ __launch_bounds__(kGB) void k_gram_finish(const double *__restrict__ partials,
                                                     int nblk, int K, int M,
                                                     double *__restrict__ C, int ldc) {
    const int e = blockIdx.x * (kGB / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (e >= K * M) return;  // the whole wave leaves together
    const double *p = partials + (long)e * nblk;
    double s = 0.0;
    for (int b = lane; b < nblk; b += 64) s += p[b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) C[(e % K) + (long)(e / K) * ldc] = s;
}

// Y = V Q: one wave per 16-row tile of Y, all J columns (JT tiles); Q staged
// in LDS; the 16 x 16 result tiles pass through LDS so that every column is
// written as 16 contiguous rows.
template <int KS, int JT>
__global__ __launch_bounds__(kGB) void k_tall_gemm_mfma(const double *__restrict__ V, long ldv,
                                                        int K, const double *__restrict__ Q,
                                                        int ldq, int J, double *__restrict__ Y,
                                                        long ldy, long n) {
    constexpr int KP = KS * 4;  // K padded to the 4-deep MFMA steps
    __shared__ double qs[KP * JT * 16];
    __shared__ double ys[4][16 * 17];
    for (int t = threadIdx.x; t < KP * JT * 16; t += kGB) {
        const int k = t % KP, j = t / KP;
        qs[t] = (k < K && j < J) ? Q[k + (long)j * ldq] : 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const long ntiles = (n + 15) / 16;
    for (long tile = (long)blockIdx.x * 4 + wave; tile < ntiles; tile += (long)gridDim.x * 4) {
        const long r0 = tile * 16;
        const long row = r0 + li;
        double a[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 4 * s + lk;
            a[s] = (row < n && k < K) ? V[(long)k * ldv + row] : 0.0;
        }
#pragma unroll
        for (int tj = 0; tj < JT; ++tj) {
            v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < KS; ++s)
                acc = mfma(a[s], qs[(4 * s + lk) + (tj * 16 + li) * KP], acc);
            // D(4*r + lk, li) -> ys[col li][row]
#pragma unroll
            for (int r = 0; r < 4; ++r) ys[wave][li * 17 + 4 * r + lk] = acc[r];
            __builtin_amdgcn_wave_barrier();
            // 4 columns per pass, 16 contiguous rows each
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const int c = 4 * p + lk;
                const int j = tj * 16 + c;
                if (j < J && r0 + li < n) Y[(long)j * ldy + r0 + li] = ys[wave][c * 17 + li];
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

}  // namespace

int gram_blocks(long n, int K, int M) {
    // partial sums of at most kGramScratch doubles: fewer blocks for larger
    // K x M, and no more blocks than 64-row steps
    long cap = kGramScratch / std::max(1, K * M);
    long steps = (n + kGR - 1) / kGR;
    long b = std::min<long>(std::min<long>(1024, cap), std::max<long>(steps, 1));
    return (int)std::max<long>(b, 1);
}

int gram(hipStream_t s, long n, int K, int M, const double *V, long ldv, const double *W,
         long ldw, double *C, int ldc, double *scratch) {
    if (K <= 0 || M <= 0) return 0;
    if (K > 64 || M > 64) return -1;
    const int nblk = gram_blocks(n, K, M);
    const int KT = (K + 15) / 16, MT = (M + 15) / 16;
#define JDHIP_GRAM(a, b) k_gram<a, b><<<nblk, kGB, 0, s>>>(V, ldv, K, W, ldw, M, n, scratch)
#define JDHIP_GRAM_M(a)              \
    do {                             \
        if (MT == 1) JDHIP_GRAM(a, 1); \
        else if (MT == 2) JDHIP_GRAM(a, 2); \
        else if (MT == 3) JDHIP_GRAM(a, 3); \
        else JDHIP_GRAM(a, 4);          \
    } while (0)
    if (KT == 1) JDHIP_GRAM_M(1);
    else if (KT == 2) JDHIP_GRAM_M(2);
    else if (KT == 3) JDHIP_GRAM_M(3);
    else JDHIP_GRAM_M(4);
#undef JDHIP_GRAM_M
#undef JDHIP_GRAM
    k_gram_finish<<<(K * M + kGB / 64 - 1) / (kGB / 64), kGB, 0, s>>>(scratch, nblk, K, M, C,
                                                                       ldc);
    return 0;
}

int tall_gemm_mfma(hipStream_t s, const double *V, long ldv, int K, const double *Q, int ldq,
                   int J, double *Y, long ldy, long n) {
    if (n <= 0 || J <= 0) return 0;
    if (K > 64 || J > 64 || K <= 0) return -1;
    const long ntiles = (n + 15) / 16;
    const int g = (int)std::max<long>(1, std::min<long>(2048, (ntiles + 3) / 4));
    const int KS = (K + 3) / 4, JT = (J + 15) / 16;
#define JDHIP_TG(ks, jt) \
    k_tall_gemm_mfma<ks, jt><<<g, kGB, 0, s>>>(V, ldv, K, Q, ldq, J, Y, ldy, n)
#define JDHIP_TG_J(ks)                \
    do {                              \
        if (JT == 1) JDHIP_TG(ks, 1);   \
        else if (JT == 2) JDHIP_TG(ks, 2); \
        else if (JT == 3) JDHIP_TG(ks, 3); \
        else JDHIP_TG(ks, 4);            \
    } while (0)
    if (KS <= 2) JDHIP_TG_J(2);
    else if (KS <= 4) JDHIP_TG_J(4);
    else if (KS <= 8) JDHIP_TG_J(8);
    else if (KS <= 12) JDHIP_TG_J(12);
    else JDHIP_TG_J(16);
#undef JDHIP_TG_J
#undef JDHIP_TG
    return 0;
}

}  // namespace jdhip
