// One 32 x 32 (op x, op y) tile of slice k of an op-frame copy (frame.hip,
// batch.hip): out[k][j][i] = img[i * sx + j * sy + k * sz], out contiguous
// [gz][gy][gx].  Lanes read along op y and write along op x through an LDS
// tile, so both sides are coalesced when op y has unit stride (the axis
// groups of a C-order image).  256 threads, every thread calls it.
#pragma once
#include "common.h"

namespace srfx {

constexpr int kFrameTile = 32;

__device__ __forceinline__ void frame_tile(const float *__restrict__ img, int gx, int gy,
                                           int64_t sx, int64_t sy, int64_t sz,
                                           float *__restrict__ out, int i0, int j0, int64_t k,
                                           float (*tile)[kFrameTile + 1])
{
    const int tx = threadIdx.x % kFrameTile, ty = threadIdx.x / kFrameTile; // 32 x 8
    for (int r = ty; r < kFrameTile; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        if (i < gx && j < gy)
            tile[r][tx] = img[i * sx + j * sy + k * sz];
    }
    __syncthreads();
    for (int r = ty; r < kFrameTile; r += 8) {
        const int i = i0 + tx, j = j0 + r;
        if (i < gx && j < gy)
            out[(k * gy + j) * gx + i] = tile[tx][r];
    }
}

} // namespace srfx
