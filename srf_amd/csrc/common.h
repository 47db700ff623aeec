// Internal helpers shared by the HIP translation units of libsrfx.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/srfx.h"

namespace srfx {

// Largest in-slice dimension whose mesh-centre table fits the kernels' LDS
// tables, and the widest supported TOR patch.
constexpr int kMaxTableDim = 1024;
constexpr int kMaxPatch = 16;
// Reference TOF switch: tof_sigma2 >= 405758 disables TOF (tor.cu.cc:20,
// siddon.cu.cc:15).
constexpr float kTofThreshold = 405758.0f;

struct Strides {
    int64_t x, y, z;
};

// Element strides of a grid; all-zero strides mean contiguous [gz][gy][gx].
inline Strides grid_strides(const srfx_grid &g)
{
    if (g.stride[0] == 0 && g.stride[1] == 0 && g.stride[2] == 0)
        return {1, (int64_t)g.shape[0], (int64_t)g.shape[0] * g.shape[1]};
    return {g.stride[0], g.stride[1], g.stride[2]};
}

inline bool finite_pos(float v) { return v > 0.0f && v < 3.0e38f; }

// Shape, size and strides sane, and every voxel offset fits an int32.
inline bool grid_ok(const srfx_grid &g)
{
    for (int k = 0; k < 3; k++) {
        if (g.shape[k] <= 0 || !finite_pos(g.size[k]))
            return false;
        if (!(g.center[k] > -3.0e38f && g.center[k] < 3.0e38f))
            return false;
    }
    Strides s = grid_strides(g);
    if (s.x < 0 || s.y < 0 || s.z < 0)
        return false;
    int64_t maxoff = (g.shape[0] - 1) * s.x + (g.shape[1] - 1) * s.y + (g.shape[2] - 1) * s.z;
    return maxoff < (int64_t)INT32_MAX;
}

inline bool same_frame(const srfx_grid &a, const srfx_grid &b)
{
    for (int k = 0; k < 3; k++)
        if (a.shape[k] != b.shape[k] || a.center[k] != b.center[k] || a.size[k] != b.size[k])
            return false;
    return true;
}

inline int64_t grid_nvox(const srfx_grid &g)
{
    return (int64_t)g.shape[0] * g.shape[1] * g.shape[2];
}

// Zero the image covered by g (strided or contiguous) unless accumulating.
int zero_grid(float *img, const srfx_grid &g, uint32_t flags, hipStream_t st);
int zero_vec(float *v, int64_t n, uint32_t flags, hipStream_t st);

inline int launch_status()
{
    return hipGetLastError() == hipSuccess ? SRFX_OK : SRFX_ELAUNCH;
}

inline unsigned blocks_for(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : b);
}

// Slab schedule field of the flags: (number of slabs, interleaved over XCDs).
inline bool slab_schedule(uint32_t flags, int &nslab, int &interleave)
{
    switch ((flags >> SRFX_SLAB_SHIFT) & 3u) {
    case 1:
        nslab = 8;
        interleave = 1;
        return true;
    case 2:
        nslab = 8;
        interleave = 0;
        return true;
    case 3:
        nslab = 16;
        interleave = 1;
        return true;
    default:
        nslab = 1;
        interleave = 0;
        return false;
    }
}

// Stream-ordered device workspace: allocated from the device's default memory
// pool (kept by the pool between calls) and released on the same stream.
int ws_alloc(void **ptr, size_t bytes, hipStream_t st);
int ws_free(void *ptr, hipStream_t st);

} // namespace srfx
