// reduce.hpp -- deterministic grid-wide reductions for gfx950 kernels.
//
// xor-butterfly inside the wave64, LDS across the waves of the block,
// per-block partials in global memory, summed in a fixed order by a second
// (finish) launch, so results do not depend on scheduling.  Accumulation is
// always in double.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace jdhip {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// This thread's share of a sum of nparts partials read by one block, in a
// fixed order.  Four independent chains: the partials come from other XCDs'
// writes (memory latency per load), so several loads stay in flight.
__device__ __forceinline__ double thread_partials(const double *__restrict__ pc, int nparts) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int b = threadIdx.x;
    for (; b + 3 * kBlock < nparts; b += 4 * kBlock) {
        s0 += pc[b];
        s1 += pc[b + kBlock];
        s2 += pc[b + 2 * kBlock];
        s3 += pc[b + 3 * kBlock];
    }
    for (; b < nparts; b += kBlock) s0 += pc[b];
    return (s0 + s1) + (s2 + s3);
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
        double s = wave_sum(thread_partials(partials + (long)c * nparts, nparts));
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
