// reduce.hpp -- deterministic grid-wide reductions for gfx950 kernels.
//
// xor-butterfly inside the wave64, LDS across the waves of the block,
// per-block partials in global memory; the last block to take a ticket from
// a device counter sums the partials in a fixed order, so results do not
// depend on scheduling.  Accumulation is always in double.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace jdhip {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Two-launch deterministic reduction.  Partial launch (nparts == 0): every
// block writes its sums to partials[c * gridDim.x + blockIdx.x] and the call
// returns false.  Finish launch (one block, nparts = number of partial blocks, the
// kernel's element loop empty): fin[c] (shared) receives the totals in a
// fixed order and the call returns true, so the caller's epilogue runs once.
// (A single-launch "last block" ticket needs one device-scope atomic per
// block on one address; across the 8 XCDs these serialise at roughly 0.1 us
// each, which dominated every reduction kernel.)
template <int NC>
__device__ bool grid_reduce(double (&v)[NC], int nc, double *__restrict__ partials, int nparts,
                            double *fin) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (nparts == 0) {
        __shared__ double sh[NC][kBlock / 64];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (c < nc) {
                double t = wave_sum(v[c]);
                if (lane == 0) sh[c][wid] = t;
            }
        }
        __syncthreads();
        if ((int)threadIdx.x < nc) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) s += sh[threadIdx.x][w];
            partials[(long)threadIdx.x * gridDim.x + blockIdx.x] = s;
        }
        return false;
    }
    __shared__ double red[kBlock / 64];
    for (int c = 0; c < nc; ++c) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nparts; b += kBlock) s += partials[(long)c * nparts + b];
        s = wave_sum(s);
        if (lane == 0) red[wid] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) t += red[w];
            fin[c] = t;
        }
        __syncthreads();
    }
    return true;
}

static inline int grid_for(long n, int per_block = kBlock, int cap = kMaxRedBlocks) {
    long b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

}  // namespace jdhip
