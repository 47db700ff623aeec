// Siddon ray-driven projector for CDNA4 (gfx950): the untiled ops, one lane
// per LOR (siddon.cu.cc:424-675).  Ray setup, the walk and its precision are
// in siddon_core.h.
//
// Unlike the reference, a ray whose start voxel falls outside the grid is
// skipped (the reference indexes out of bounds there), and the walk is capped
// at gx+gy+gz+3 steps; both only affect degenerate rays.
#include "siddon_core.h"

namespace srfx {
namespace {

template <int MODE>
__global__ __launch_bounds__(256) void siddon_kernel(SidParams p)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int cnt = 0;
    if (e < p.n) {
        Walk w;
        setup_walk<MODE>(p, e, w);
        const bool tof = (MODE != kMap) && p.use_tof;
        if (MODE == kProj) {
            float pv = 0.0f;
            if (w.ok)
                pv = tof ? walk<0, true>(p, w, 0.0f, nullptr) : walk<0, false>(p, w, 0.0f, nullptr);
            p.out[e] = pv;
        } else if (MODE == kBack || MODE == kMap) {
            float v = p.val[e];
            if (w.ok && v > 0.0f) {
                if (tof)
                    walk<1, true>(p, w, 1.0f / v, nullptr);
                else
                    walk<1, false>(p, w, 1.0f / v, nullptr);
            }
        } else if (MODE == kAdj) {
            float g = p.val[e];
            if (w.ok && g != 0.0f) {
                if (tof)
                    walk<1, true>(p, w, g, nullptr);
                else
                    walk<1, false>(p, w, g, nullptr);
            }
        } else if (MODE == kOsem) {
            if (w.ok) {
                float pv = tof ? walk<0, true>(p, w, 0.0f, nullptr)
                               : walk<0, false>(p, w, 0.0f, nullptr);
                if (pv > 0.0f) {
                    if (tof)
                        walk<1, true>(p, w, 1.0f / pv, nullptr);
                    else
                        walk<1, false>(p, w, 1.0f / pv, nullptr);
                }
            }
        } else { // kCount
            if (w.ok) {
                if (tof)
                    walk<2, true>(p, w, 0.0f, &cnt);
                else
                    walk<2, false>(p, w, 0.0f, &cnt);
            }
            if (p.nnz)
                p.nnz[e] = cnt;
        }
    }
    if (MODE == kCount && p.nnz_total) {
        unsigned long long c = (unsigned long long)cnt;
        for (int off = 32; off > 0; off >>= 1)
            c += __shfl_down(c, off, 64);
        if ((threadIdx.x & 63) == 0 && c)
            atomicAdd(p.nnz_total, c);
    }
}

int setup(SidParams &p, const float *lors, int64_t n, int64_t ld, const srfx_grid &g,
          float tof_bin, float tof_sigma2)
{
    if (n < 0 || (n > 0 && (!lors || ld < n)) || !grid_ok(g))
        return SRFX_EINVAL;
    p = SidParams{};
    p.lors = lors;
    p.n = n;
    p.ld = ld;
    for (int c = 0; c < 3; c++) {
        p.g[c] = g.shape[c];
        p.size[c] = g.size[c];
        p.center[c] = g.center[c];
    }
    Strides s = grid_strides(g);
    p.s[0] = p.o[0] = (int)s.x;
    p.s[1] = p.o[1] = (int)s.y;
    p.s[2] = p.o[2] = (int)s.z;
    p.tof_bin = tof_bin;
    p.tof_sigma2 = tof_sigma2;
    p.use_tof = tof_sigma2 < kTofThreshold ? 1 : 0;
    return SRFX_OK;
}

template <int MODE>
int launch(const SidParams &p, hipStream_t st)
{
    if (p.n == 0)
        return SRFX_OK;
    hipLaunchKernelGGL(siddon_kernel<MODE>, dim3(blocks_for(p.n, 256)), dim3(256), 0, st, p);
    return launch_status();
}

} // namespace
} // namespace srfx

using namespace srfx;

extern "C" int srfx_siddon_project(const float *lors, int64_t n, int64_t ld, const float *image,
                                   srfx_grid grid, float tof_bin, float tof_sigma2, float *proj,
                                   uint32_t flags, void *stream)
{
    (void)flags;
    SidParams p;
    int rc = setup(p, lors, n, ld, grid, tof_bin, tof_sigma2);
    if (rc)
        return rc;
    if (n > 0 && (!image || !proj))
        return SRFX_EINVAL;
    p.img = image;
    p.out = proj;
    return launch<kProj>(p, (hipStream_t)stream);
}

static int sid_scatter(int mode, const float *lors, int64_t n, int64_t ld, const float *vals,
                       srfx_grid grid, float tof_bin, float tof_sigma2, float *image,
                       uint32_t flags, void *stream)
{
    SidParams p;
    int rc = setup(p, lors, n, ld, grid, tof_bin, tof_sigma2);
    if (rc)
        return rc;
    if (!image || (n > 0 && !vals))
        return SRFX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    rc = zero_grid(image, grid, flags, st);
    if (rc)
        return rc;
    p.val = vals;
    p.out = image;
    switch (mode) {
    case kBack:
        return launch<kBack>(p, st);
    case kAdj:
        return launch<kAdj>(p, st);
    default:
        return launch<kMap>(p, st);
    }
}

extern "C" int srfx_siddon_backproject(const float *lors, int64_t n, int64_t ld,
                                       const float *lors_value, srfx_grid grid, float tof_bin,
                                       float tof_sigma2, float *image, uint32_t flags,
                                       void *stream)
{
    return sid_scatter(kBack, lors, n, ld, lors_value, grid, tof_bin, tof_sigma2, image, flags,
                       stream);
}

extern "C" int srfx_siddon_backproject_adjoint(const float *lors, int64_t n, int64_t ld,
                                               const float *g, srfx_grid grid, float tof_bin,
                                               float tof_sigma2, float *image, uint32_t flags,
                                               void *stream)
{
    return sid_scatter(kAdj, lors, n, ld, g, grid, tof_bin, tof_sigma2, image, flags, stream);
}

extern "C" int srfx_siddon_maplors(const float *lors, int64_t n, int64_t ld,
                                   const float *lors_value, srfx_grid grid, float *image,
                                   uint32_t flags, void *stream)
{
    return sid_scatter(kMap, lors, n, ld, lors_value, grid, 0.0f, 1.0e6f, image, flags, stream);
}

extern "C" int srfx_siddon_osem_subset(const float *lors, int64_t n, int64_t ld,
                                       const float *image, srfx_grid grid, float tof_bin,
                                       float tof_sigma2, float *bp, uint32_t flags, void *stream)
{
    SidParams p;
    int rc = setup(p, lors, n, ld, grid, tof_bin, tof_sigma2);
    if (rc)
        return rc;
    if (!bp || (n > 0 && !image))
        return SRFX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    rc = zero_grid(bp, grid, flags, st);
    if (rc)
        return rc;
    p.img = image;
    p.out = bp;
    return launch<kOsem>(p, st);
}

extern "C" int srfx_siddon_count_nnz(const float *lors, int64_t n, int64_t ld, srfx_grid grid,
                                     float tof_bin, float tof_sigma2, int32_t *nnz_per_event,
                                     uint64_t *nnz_total, void *stream)
{
    SidParams p;
    int rc = setup(p, lors, n, ld, grid, tof_bin, tof_sigma2);
    if (rc)
        return rc;
    p.nnz = nnz_per_event;
    p.nnz_total = (unsigned long long *)nnz_total;
    return launch<kCount>(p, (hipStream_t)stream);
}
