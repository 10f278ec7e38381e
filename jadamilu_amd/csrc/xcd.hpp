// xcd.hpp -- XCD-aware block numbering for gfx950 kernels (HIP sources only).
//
// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8), each with its own L2.  logical_block() renumbers the blocks so
// that every XCD works on one contiguous range of the grid: neighbouring rows
// of a sparse product share x entries, and the shared cache lines then stay
// in one L2 instead of being fetched by all eight.  The map is a bijection of
// 0 .. gridDim.x-1 for every grid size.
#pragma once
#include <hip/hip_runtime.h>

namespace jdhip {

__device__ __forceinline__ long logical_block(int xcd) {
    if (!xcd) return blockIdx.x;
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned per = nb / 8, rem = nb % 8, x = b % 8, i = b / 8;
    return (long)x * per + (x < rem ? x : rem) + i;
}

}  // namespace jdhip
