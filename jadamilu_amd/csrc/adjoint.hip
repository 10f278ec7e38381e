// adjoint.hip -- eigenvalue gradients on the sparsity pattern (gfx950).
//
// For a simple eigenvalue lambda_k of the symmetric A with unit eigenvector
// x_k, d lambda_k / d A_ij = x_k(i) x_k(j).  The loss sum_k w_k lambda_k
// therefore has, on the stored upper-triangle CSR (each off-diagonal entry
// stands for A_ij and A_ji), the gradient
//     g[e(i,j)] = c_ij * sum_k w_k x_k(i) x_k(j),   c_ij = 2 (i != j), 1 (i == j),
// and for a diagonal potential A = H + diag(V):  g_V[i] = sum_k w_k x_k(i)^2.
// Valid while each lambda_k is separated from the rest of the spectrum.
// No reference counterpart (the reference has no derivatives).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/jadamilu_amd.h"

namespace jdhip {

constexpr int kAB = 256;
constexpr int kMaxK = 64;

struct Weights {
    double w[kMaxK];
};

// one group of 4 lanes per row of the 1-based upper CSR
__global__ __launch_bounds__(kAB) void k_adjoint_csr(int n, const int *__restrict__ ia,
                                                     const int *__restrict__ ja,
                                                     const double *__restrict__ X, long ldx,
                                                     int k, Weights w, double *__restrict__ g) {
    const int grp = threadIdx.x >> 2, l = threadIdx.x & 3;
    for (long i = (long)blockIdx.x * (kAB / 4) + grp; i < n; i += (long)gridDim.x * (kAB / 4)) {
        const int b = ia[i] - 1, e = ia[i + 1] - 1;
        for (int t = b + l; t < e; t += 4) {
            const int j = ja[t] - 1;
            double s = 0.0;
            for (int c = 0; c < k; ++c) s += w.w[c] * X[c * ldx + i] * X[c * ldx + j];
            g[t] = (j == i) ? s : 2.0 * s;
        }
    }
}

__global__ __launch_bounds__(kAB) void k_adjoint_diag(int n, const double *__restrict__ X,
                                                      long ldx, int k, Weights w,
                                                      double *__restrict__ g) {
    for (long i = blockIdx.x * (long)kAB + threadIdx.x; i < n; i += (long)gridDim.x * kAB) {
        double s = 0.0;
        for (int c = 0; c < k; ++c) {
            double v = X[c * ldx + i];
            s += w.w[c] * v * v;
        }
        g[i] = s;
    }
}

}  // namespace jdhip

extern "C" {

int jdhip_adjoint_csr(void *stream, int n, const int *ia_dev, const int *ja_dev,
                      const double *x_dev, long ldx, int k, const double *w, double *g_dev) {
    if (k <= 0 || k > jdhip::kMaxK || n <= 0) return -1;
    jdhip::Weights W{};
    for (int c = 0; c < k; ++c) W.w[c] = w[c];
    int blocks = (int)std::min<long>(8192, ((long)n + jdhip::kAB / 4 - 1) / (jdhip::kAB / 4));
    jdhip::k_adjoint_csr<<<std::max(blocks, 1), jdhip::kAB, 0, (hipStream_t)stream>>>(
        n, ia_dev, ja_dev, x_dev, ldx, k, W, g_dev);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int jdhip_adjoint_diag(void *stream, int n, const double *x_dev, long ldx, int k,
                       const double *w, double *g_dev) {
    if (k <= 0 || k > jdhip::kMaxK || n <= 0) return -1;
    jdhip::Weights W{};
    for (int c = 0; c < k; ++c) W.w[c] = w[c];
    int blocks = (int)std::min<long>(8192, ((long)n + jdhip::kAB - 1) / jdhip::kAB);
    jdhip::k_adjoint_diag<<<std::max(blocks, 1), jdhip::kAB, 0, (hipStream_t)stream>>>(
        n, x_dev, ldx, k, W, g_dev);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
