// rowsum.hpp -- the row sum of a CSR product by a lane group (HIP sources
// only), shared by the SpMV and the factor's CSR kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace jdhip {

// Sum of val[k] * x[idx[k]] over the entries b + l, b + l + LPR, ... (< e)
// of one CSR row, in that order: one lane of a group of LPR.  Four entries
// are loaded at a time, all four indices first, then the values and the
// gathers, with no branch between the loads (positions past the row's end
// load its last entry again and are not added), so a lane keeps up to twelve
// loads in flight.
template <int LPR, typename T, typename W>
__device__ __forceinline__ double csr_lane_sum(const int *__restrict__ idx,
                                               const T *__restrict__ val, const W *x, int b,
                                               int e, int l) {
    double s = 0.0;
    for (int k0 = b + l; k0 < e; k0 += 4 * LPR) {
        int c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = idx[min(k0 + u * LPR, e - 1)];
        double p[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) p[u] = (double)val[min(k0 + u * LPR, e - 1)] * (double)x[c[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k0 + u * LPR < e) s += p[u];
    }
    return s;
}

}  // namespace jdhip
