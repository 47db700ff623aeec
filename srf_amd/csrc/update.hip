// Elementwise kernels: output zeroing (a fill kernel, so the ops stay
// capturable in HIP graphs) and the MLEM image update.
#include <atomic>

#include "common.h"

namespace srfx {
namespace {

__global__ __launch_bounds__(256) void fill_contig(float *v, int64_t n, float value)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride)
        v[i] = value;
}

__global__ __launch_bounds__(256) void fill_strided(float *v, int gx, int gy, int gz, int sx,
                                                    int sy, int sz, float value)
{
    int64_t n = (int64_t)gx * gy * gz;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        int ix = (int)(i % gx);
        int iy = (int)((i / gx) % gy);
        int iz = (int)(i / ((int64_t)gx * gy));
        v[ix * sx + iy * sy + iz * sz] = value;
    }
}

// model/recon_step.py:48-49: d * efficiency_map * image_succ.
__global__ __launch_bounds__(256) void mlem_update_kernel(float *x, const float *__restrict__ e,
                                                          const float *__restrict__ b, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride)
        x[i] = (x[i] * e[i]) * b[i];
}

unsigned fill_blocks(int64_t n)
{
    int64_t b = (n + 255) / 256;
    if (b > 8192)
        b = 8192;
    return (unsigned)(b < 1 ? 1 : b);
}

} // namespace

int zero_vec(float *v, int64_t n, uint32_t flags, hipStream_t st)
{
    if (flags & SRFX_ACCUMULATE)
        return SRFX_OK;
    if (n <= 0)
        return SRFX_OK;
    hipLaunchKernelGGL(fill_contig, dim3(fill_blocks(n)), dim3(256), 0, st, v, n, 0.0f);
    return launch_status();
}

int zero_grid(float *img, const srfx_grid &g, uint32_t flags, hipStream_t st)
{
    if (flags & SRFX_ACCUMULATE)
        return SRFX_OK;
    Strides s = grid_strides(g);
    int64_t n = grid_nvox(g);
    if (s.x == 1 && s.y == g.shape[0] && s.z == (int64_t)g.shape[0] * g.shape[1])
        return zero_vec(img, n, flags, st);
    hipLaunchKernelGGL(fill_strided, dim3(fill_blocks(n)), dim3(256), 0, st, img, g.shape[0],
                       g.shape[1], g.shape[2], (int)s.x, (int)s.y, (int)s.z, 0.0f);
    return launch_status();
}

int ws_alloc(void **ptr, size_t bytes, hipStream_t st)
{
    // Keep freed workspace in the pool instead of returning it to the driver
    // at every synchronisation (once per device).
    static std::atomic<uint64_t> pool_set{0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return SRFX_ELAUNCH;
    uint64_t bit = 1ull << (dev & 63);
    if (!(pool_set.load() & bit)) {
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t thr = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
        pool_set.fetch_or(bit);
    }
    if (hipMallocAsync(ptr, bytes ? bytes : 256, st) != hipSuccess) {
        (void)hipGetLastError();
        return SRFX_ENOMEM;
    }
    return SRFX_OK;
}

int ws_free(void *ptr, hipStream_t st)
{
    return hipFreeAsync(ptr, st) == hipSuccess ? SRFX_OK : SRFX_ELAUNCH;
}

} // namespace srfx

using namespace srfx;

extern "C" int srfx_mlem_update(float *x, const float *emap, const float *bp, int64_t n,
                                void *stream)
{
    if (n < 0 || (n > 0 && (!x || !emap || !bp)))
        return SRFX_EINVAL;
    if (n == 0)
        return SRFX_OK;
    hipLaunchKernelGGL(mlem_update_kernel, dim3(fill_blocks(n)), dim3(256), 0,
                       (hipStream_t)stream, x, emap, bp, n);
    return launch_status();
}

extern "C" const char *srfx_strerror(int status)
{
    switch (status) {
    case SRFX_OK:
        return "ok";
    case SRFX_EINVAL:
        return "invalid argument";
    case SRFX_ELAUNCH:
        return "HIP launch failed";
    case SRFX_EUNSUPPORTED:
        return "unsupported configuration";
    case SRFX_ENOMEM:
        return "device workspace allocation failed";
    case SRFX_ETOOLARGE:
        return "input larger than the entry point's static scratch";
    default:
        return "unknown status";
    }
}
