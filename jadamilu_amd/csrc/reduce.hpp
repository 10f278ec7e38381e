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

// Returns true in the last block, where fin[c] (shared) holds the totals.
template <int NC>
__device__ bool grid_reduce(double (&v)[NC], int nc, double *__restrict__ partials,
                            unsigned *__restrict__ counter, double *fin) {
    __shared__ double sh[NC][kBlock / 64];
    __shared__ int amlast;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
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
        partials[(long)blockIdx.x * NC + threadIdx.x] = s;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) amlast = (atomicAdd(counter, 1u) == gridDim.x - 1);
    __syncthreads();
    if (!amlast) return false;
    __threadfence();
    for (int c = wid; c < nc; c += kBlock / 64) {
        double s = 0.0;
        for (int b = lane; b < (int)gridDim.x; b += 64) s += partials[(long)b * NC + c];
        s = wave_sum(s);
        if (lane == 0) fin[c] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) *counter = 0u;
    return true;
}

static inline int grid_for(long n, int per_block = kBlock, int cap = kMaxRedBlocks) {
    long b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

}  // namespace jdhip
