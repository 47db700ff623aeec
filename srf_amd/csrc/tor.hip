// TOF tube-of-response (TOR) projector entry points (C ABI, include/srfx.h).
// The device code is in tor_core.h.
#include <mutex>

#include "tor_core.h"

using namespace srfx;

// Slab projection: per-slab partial sums in a workspace, then their sum in
// slab order into p.out.
static int slab_project(const TorParams &p, int nslab, uint32_t flags, hipStream_t st)
{
    void *ws = nullptr;
    int rc = ws_alloc(&ws, sizeof(float) * (size_t)nslab * (size_t)p.n, st);
    if (rc)
        return rc;
    rc = launch_slab<kProj>(p, (float *)ws, flags, st);
    if (!rc)
        rc = launch_slab_sum((const float *)ws, nslab, p.n, p.out, st);
    int rf = ws_free(ws, st);
    return rc ? rc : rf;
}

extern "C" int srfx_tor_project(const float *lors, int64_t n, int64_t ld, const float *image,
                                srfx_grid grid, srfx_tor_model m, float *proj, uint32_t flags,
                                void *stream)
{
    TorParams p;
    int rc = setup(p, lors, n, ld, grid, m.kernel_width, m.geo_sigma2_flag, m.tof_bin,
                   m.tof_sigma2);
    if (rc)
        return rc;
    if (n > 0 && (!image || !proj))
        return SRFX_EINVAL;
    p.full_patch = (flags & SRFX_FULL_PATCH) ? 1 : 0;
    p.img = image;
    p.out = proj;
    hipStream_t st = (hipStream_t)stream;
    int nslab, inter;
    if (!slab_schedule(flags, nslab, inter) || n == 0)
        return launch<kProj>(p, flags, st);
    return slab_project(p, nslab, flags, st);
}

static int tor_scatter_common(int mode, const float *lors, int64_t n, int64_t ld,
                              const float *vals, srfx_grid grid, srfx_tor_model m, float *image,
                              uint32_t flags, void *stream)
{
    TorParams p;
    int rc = setup(p, lors, n, ld, grid, m.kernel_width, m.geo_sigma2_flag, m.tof_bin,
                   m.tof_sigma2);
    if (rc)
        return rc;
    if (!image || (n > 0 && !vals))
        return SRFX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    rc = zero_grid(image, grid, flags, st);
    if (rc)
        return rc;
    p.full_patch = (flags & SRFX_FULL_PATCH) ? 1 : 0;
    p.val = vals;
    p.out = image;
    int nslab, inter;
    if (slab_schedule(flags, nslab, inter)) {
        switch (mode) {
        case kBack:
            return launch_slab<kBack>(p, nullptr, flags, st);
        case kAdj:
            return launch_slab<kAdj>(p, nullptr, flags, st);
        default:
            return launch_slab<kMap>(p, nullptr, flags, st);
        }
    }
    switch (mode) {
    case kBack:
        return launch<kBack>(p, flags, st);
    case kAdj:
        return launch<kAdj>(p, flags, st);
    default:
        return launch<kMap>(p, flags, st);
    }
}

extern "C" int srfx_tor_backproject(const float *lors, int64_t n, int64_t ld,
                                    const float *lors_value, srfx_grid grid, srfx_tor_model m,
                                    float *image, uint32_t flags, void *stream)
{
    return tor_scatter_common(kBack, lors, n, ld, lors_value, grid, m, image, flags, stream);
}

extern "C" int srfx_tor_backproject_adjoint(const float *lors, int64_t n, int64_t ld,
                                            const float *g, srfx_grid grid, srfx_tor_model m,
                                            float *image, uint32_t flags, void *stream)
{
    return tor_scatter_common(kAdj, lors, n, ld, g, grid, m, image, flags, stream);
}

extern "C" int srfx_tor_maplors(const float *lors, int64_t n, int64_t ld, const float *lors_value,
                                srfx_grid grid, float kernel_width, int32_t geo_sigma2_flag,
                                float *image, uint32_t flags, void *stream)
{
    return tor_scatter_common(kMap, lors, n, ld, lors_value, grid,
                              model_no_tof(kernel_width, geo_sigma2_flag), image, flags, stream);
}

// Arguments of srfx_tor_osem_subset checked and set up, bp zeroed unless
// accumulating.  nslab = 0 when the flags name no slab schedule.
static int osem_subset_setup(TorParams &p, const float *lors, int64_t n, int64_t ld,
                             const float *image, const srfx_grid &image_grid,
                             const srfx_grid &bp_grid, const srfx_tor_model &m, float *bp,
                             uint32_t flags, hipStream_t st, int &nslab)
{
    int rc = setup(p, lors, n, ld, image_grid, m.kernel_width, m.geo_sigma2_flag, m.tof_bin,
                   m.tof_sigma2);
    if (rc)
        return rc;
    if (!grid_ok(bp_grid) || !same_frame(image_grid, bp_grid) || !bp || (n > 0 && !image))
        return SRFX_EINVAL;
    int inter;
    if (!slab_schedule(flags, nslab, inter))
        nslab = 0;
    rc = zero_grid(bp, bp_grid, flags, st);
    if (rc)
        return rc;
    Strides so = grid_strides(bp_grid);
    p.ox = (int)so.x;
    p.oy = (int)so.y;
    p.oz = (int)so.z;
    p.full_patch = (flags & SRFX_FULL_PATCH) ? 1 : 0;
    p.img = image;
    p.out = bp;
    return SRFX_OK;
}

// Slab schedule of the OSEM pass: the per-slab projections into part (nslab
// x n floats, every one written before it is read), then the scatter with
// their sum.
static int osem_subset_slabs(const TorParams &p, float *part, uint32_t flags, hipStream_t st)
{
    int rc = launch_slab<kProj>(p, part, flags, st);
    return rc ? rc : launch_slab<kOsem>(p, part, flags, st);
}

extern "C" int srfx_tor_osem_subset(const float *lors, int64_t n, int64_t ld, const float *image,
                                    srfx_grid image_grid, srfx_grid bp_grid, srfx_tor_model m,
                                    float *bp, uint32_t flags, void *stream)
{
    TorParams p;
    hipStream_t st = (hipStream_t)stream;
    int nslab = 0;
    int rc = osem_subset_setup(p, lors, n, ld, image, image_grid, bp_grid, m, bp, flags, st,
                               nslab);
    if (rc)
        return rc;
    if (nslab == 0 || n == 0)
        return launch<kOsem>(p, flags, st);
    void *ws = nullptr;
    rc = ws_alloc(&ws, sizeof(float) * (size_t)nslab * (size_t)n, st);
    if (rc)
        return rc;
    rc = osem_subset_slabs(p, (float *)ws, flags, st);
    int rf = ws_free(ws, st);
    return rc ? rc : rf;
}

// Scratch of srfx_tor_osem_subset_noalloc: the per-slab partial projections,
// one copy per device (a module global), 16 MiB.  Host code never passes the
// array's name to a kernel (on the host it is only a handle of the symbol);
// it passes the device address hipGetSymbolAddress returns, looked up once
// per device on the entry point's first call there.
constexpr int64_t kScratchFloats = (int64_t)1 << 22;
__device__ __align__(256) float scratch[kScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
float *scratch_dev[kMaxDevices];

// Device address of the scratch on the current device.
int scratch_address(float **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess)
        return SRFX_ELAUNCH;
    if (dev < 0 || dev >= kMaxDevices)
        return SRFX_EUNSUPPORTED;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return SRFX_ELAUNCH;
        }
        scratch_dev[dev] = (float *)ptr;
    }
    *out = scratch_dev[dev];
    return SRFX_OK;
}
} // namespace

extern "C" int srfx_tor_osem_subset_noalloc(const float *lors, int64_t n, int64_t ld,
                                            const float *image, srfx_grid image_grid,
                                            srfx_grid bp_grid, srfx_tor_model m, float *bp,
                                            uint32_t flags, void *stream)
{
    // the slab schedule of the flags, SRFX_SLAB_XCD8 when they name none
    if (((flags >> SRFX_SLAB_SHIFT) & 3u) == 0)
        flags |= SRFX_SLAB_XCD8;
    int nslab = 0, inter = 0;
    (void)slab_schedule(flags, nslab, inter);
    // size checks before any device work
    if (n > 0 && (int64_t)nslab * n > kScratchFloats)
        return SRFX_ETOOLARGE;
    float *part = nullptr;
    int rc = scratch_address(&part);
    if (rc)
        return rc;
    TorParams p;
    hipStream_t st = (hipStream_t)stream;
    rc = osem_subset_setup(p, lors, n, ld, image, image_grid, bp_grid, m, bp, flags, st, nslab);
    if (rc || n == 0)
        return rc;
    return osem_subset_slabs(p, part, flags, st);
}

extern "C" int srfx_tor_count_nnz(const float *lors, int64_t n, int64_t ld, srfx_grid grid,
                                  srfx_tor_model m, int32_t *nnz_per_event, uint64_t *nnz_total,
                                  void *stream)
{
    TorParams p;
    int rc = setup(p, lors, n, ld, grid, m.kernel_width, m.geo_sigma2_flag, 0.0f, 1.0e6f);
    if (rc)
        return rc;
    p.nnz = nnz_per_event;
    p.nnz_total = (unsigned long long *)nnz_total;
    p.full_patch = 0;
    return launch<kCount>(p, 0u, (hipStream_t)stream);
}

extern "C" int srfx_tor_patch_info(srfx_grid grid, float kernel_width, int32_t *patch,
                                   int32_t *box)
{
    TorParams p;
    const float one = 1.0f;
    int rc = setup(p, &one, 0, 0, grid, kernel_width, 0, 0.0f, 1.0e6f);
    if (rc)
        return rc;
    if (patch)
        *patch = p.patch;
    if (box)
        *box = p.box;
    return SRFX_OK;
}
