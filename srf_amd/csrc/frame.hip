// Copy of an image into an axis group's op frame: out[k][j][i] =
// img[i * sx + j * sy + k * sz], out contiguous [gz][gy][gx].  This is the
// reference's transpose into the group frame (split_lors.py:79, perm(a)) done
// as one pass: for the axis groups of a C-order (nx, ny, nz) image the op y
// axis is world z (unit stride) and op x has a large stride, so each
// workgroup moves a 32 x 32 (op x, op y) tile of one slice through LDS,
// reading along op y and writing along op x, both coalesced.
#include "frame_core.h"

namespace srfx {
namespace {

__global__ __launch_bounds__(256) void frame_copy_kernel(const float *__restrict__ img, int gx,
                                                         int gy, int64_t sx, int64_t sy,
                                                         int64_t sz, float *__restrict__ out)
{
    __shared__ float tile[kFrameTile][kFrameTile + 1];
    frame_tile(img, gx, gy, sx, sy, sz, out, blockIdx.x * kFrameTile, blockIdx.y * kFrameTile,
               blockIdx.z, tile);
}

} // namespace
} // namespace srfx

extern "C" int srfx_frame_copy(const float *img, srfx_grid grid, float *out, void *stream)
{
    using namespace srfx;
    if (!img || !out || img == out || !grid_ok(grid))
        return SRFX_EINVAL;
    const Strides s = grid_strides(grid);
    const dim3 blocks((grid.shape[0] + kFrameTile - 1) / kFrameTile,
                      (grid.shape[1] + kFrameTile - 1) / kFrameTile, grid.shape[2]);
    hipLaunchKernelGGL(frame_copy_kernel, blocks, dim3(256), 0, (hipStream_t)stream, img,
                       grid.shape[0], grid.shape[1], s.x, s.y, s.z, out);
    return launch_status();
}
