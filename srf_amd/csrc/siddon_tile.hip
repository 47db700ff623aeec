// Tiled Siddon backprojection over a fixed LOR set (srfx_siddon_plan_*).
//
// The untiled scatter (siddon.hip) adds every segment of every ray to the
// image with its own global FP32 atomic.  Those atomics land on scattered
// cache lines and bound the op: one C1 subset (250k LORs, 107 segments each)
// is 27M single-lane atomic requests.  A plan splits each ray into runs, the
// maximal stretches of its walk inside one image tile, and lists the runs by
// tile.  A workgroup then takes up to `chunk` runs of one tile, accumulates
// them in an FP64 copy of the tile in LDS and flushes the tile once with
// coalesced atomics.
//
// Each run's listing holds the walk state at the run's first segment (voxel,
// nextT per axis, the previous crossing) as the plan's walk, which is the
// reference's RayTracing walk (siddon.cu.cc:293-331, siddon_core.h), reached
// it, and the number of segments of the run.  The scatter resumes the walk
// there with the same FP64 operations, so every run visits the voxels and
// forms the segment values (seg_value) of the untiled walk bit for bit; only
// the order of the sums changes (FP64 within a tile).
//
// Layout.  Ray record (3 float4 per ray): [deltaT0 deltaT1] [deltaT2 t_tof]
// (FP64 pairs), [weight, steps, -, -]; listing (3 float4 per run):
// [ray, x | y << 16, z | len << 16, -], [nextT0 nextT1], [nextT2 last_t].
#include <algorithm>
#include <atomic>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "siddon_core.h"

namespace srfx {
namespace {

constexpr int kRayF4 = 3;
constexpr int kRunF4 = 3;
// LDS tile: FP64, at most 64 KB.
constexpr int kMaxTileVox = 8192;

struct SidTileGeom {
    int t[3];   // tile extent per image axis (voxels)
    int nt[3];  // tiles per axis
    int ls[3];  // LDS stride of each axis inside a tile
    int ord[3]; // axes from the fastest LDS axis to the slowest
    int tvox;   // voxels per tile
};

struct SidChunk {
    int64_t start;
    int tile, count;
};

__device__ __forceinline__ float4 pack_d2(double a, double b)
{
    return make_float4(__int_as_float(__double2loint(a)), __int_as_float(__double2hiint(a)),
                       __int_as_float(__double2loint(b)), __int_as_float(__double2hiint(b)));
}

__device__ __forceinline__ void unpack_d2(const float4 &f, double &a, double &b)
{
    a = __hiloint2double(__float_as_int(f.y), __float_as_int(f.x));
    b = __hiloint2double(__float_as_int(f.w), __float_as_int(f.z));
}

__device__ __forceinline__ int tile_id(const SidTileGeom &g, int x, int y, int z)
{
    return ((x / g.t[0]) * g.nt[1] + y / g.t[1]) * g.nt[2] + z / g.t[2];
}

// Steps along an axis until the walk leaves its current tile on that axis.
__device__ __forceinline__ int tile_steps(int pos, int step, int t)
{
    return step > 0 ? t - pos % t : (step < 0 ? pos % t + 1 : 0x7fffffff);
}

// Walk a ray as siddon_core.h's walk() does (same argmin, same FP64
// updates, same end conditions), cut into runs at tile boundaries, and call
// f(tile, x, y, z, len, nextT0, nextT1, nextT2, last_t) with each run's first
// voxel and walk state for every run with at least one kept segment (all of
// them without TOF; t^2 / sigma^2 < 9 with TOF).
template <bool TOF, typename F>
__device__ __forceinline__ void for_each_run(const SidParams &p, const SidTileGeom &g,
                                             const Walk &w, F &&f)
{
    int x = w.pos[0], y = w.pos[1], z = w.pos[2];
    double n0 = w.nextT[0], n1 = w.nextT[1], n2 = w.nextT[2];
    const double t0 = w.deltaT[0], t1 = w.deltaT[1], t2 = w.deltaT[2];
    int b0 = w.inBuf[0], b1 = w.inBuf[1], b2 = w.inBuf[2];
    int r0 = tile_steps(x, w.step[0], g.t[0]);
    int r1 = tile_steps(y, w.step[1], g.t[1]);
    int r2 = tile_steps(z, w.step[2], g.t[2]);
    double last_t = 0.0;
    int stillin = b0 * b1 * b2;
    int cap = p.g[0] + p.g[1] + p.g[2] + 3;
    // the current run: its first voxel and walk state, length, kept segments
    int sx = x, sy = y, sz = z, len = 0;
    double s0 = n0, s1 = n1, s2 = n2, slast = 0.0;
    bool any = false;
    while (stillin && cap-- > 0) {
        int i = (n0 <= n1) ? ((n0 <= n2) ? 0 : 2) : ((n1 <= n2) ? 1 : 2);
        double nt = i == 0 ? n0 : (i == 1 ? n1 : n2);
        float gv;
        any |= seg_value<TOF>(p, w.weight, w.t_tof, nt, last_t, gv);
        len++;
        last_t = nt;
        bool cross;
        if (i == 0) {
            n0 = __dadd_rn(nt, t0);
            x += w.step[0];
            stillin = --b0;
            cross = --r0 == 0;
            r0 = cross ? g.t[0] : r0;
        } else if (i == 1) {
            n1 = __dadd_rn(nt, t1);
            y += w.step[1];
            stillin = --b1;
            cross = --r1 == 0;
            r1 = cross ? g.t[1] : r1;
        } else {
            n2 = __dadd_rn(nt, t2);
            z += w.step[2];
            stillin = --b2;
            cross = --r2 == 0;
            r2 = cross ? g.t[2] : r2;
        }
        if (cross) {
            if (any)
                f(tile_id(g, sx, sy, sz), sx, sy, sz, len, s0, s1, s2, slast);
            sx = x;
            sy = y;
            sz = z;
            s0 = n0;
            s1 = n1;
            s2 = n2;
            slast = last_t;
            len = 0;
            any = false;
        }
    }
    if (len > 0 && any)
        f(tile_id(g, sx, sy, sz), sx, sy, sz, len, s0, s1, s2, slast);
}

__device__ __forceinline__ void put_run(float4 *o, int el, int sx, int sy, int sz, int len,
                                        double s0, double s1, double s2, double slast)
{
    o[0] = make_float4(__int_as_float(el), __int_as_float(sx | (sy << 16)),
                       __int_as_float(sz | (len << 16)), 0.0f);
    o[1] = pack_d2(s0, s1);
    o[2] = pack_d2(s2, slast);
}

// The plan set's two passes over the rays of all its plans (plan k holds
// rays bounds[k] .. bounds[k + 1] - 1), one lane per ray.  Count pass (FILL
// false): writes the ray records and counts the listed runs of each
// (plan, tile) counter.  Fill pass: writes each run at its counter's cursor
// (which starts at the counter's offset), with the ray's index inside its
// plan.  LDS: the workgroup first counts its runs in LDS counters, so that
// the global counters take one atomic per (workgroup, counter) instead of
// one per run; the fill pass then reserves one range per (workgroup,
// counter) and walks its rays a second time to write into those ranges.
// Needs 12 bytes of LDS per counter.
template <bool FILL, bool TOF, bool LDS>
__global__ __launch_bounds__(256) void sid_plan_pass(SidParams p, SidTileGeom g,
                                                     unsigned long long *tiles, float4 *rays,
                                                     float4 *runs, unsigned long long nruns,
                                                     const int64_t *__restrict__ bounds,
                                                     int nplans, int ntiles)
{
    extern __shared__ unsigned lcnt[];
    const int ncnt = nplans * ntiles;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = e < p.n && e >= bounds[0] && e < bounds[nplans];
    // the plan of ray e: the last k with bounds[k] <= e
    int lo = 0;
    if (valid) {
        int hi = nplans - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (bounds[mid] <= e)
                lo = mid;
            else
                hi = mid - 1;
        }
    }
    const int c0 = lo * ntiles;
    const int el = valid ? (int)(e - bounds[lo]) : 0;
    Walk w;
    w.ok = false;
    if (valid) {
        setup_walk<kBack>(p, e, w);
        if (!FILL) {
            float4 *r = rays + kRayF4 * e;
            r[0] = pack_d2(w.ok ? w.deltaT[0] : 0.0, w.ok ? w.deltaT[1] : 0.0);
            r[1] = pack_d2(w.ok ? w.deltaT[2] : 0.0, w.ok ? w.t_tof : 0.0);
            const int steps =
                w.ok ? (w.step[0] + 1) | ((w.step[1] + 1) << 2) | ((w.step[2] + 1) << 4) : 0;
            r[2] = make_float4(w.weight, __int_as_float(steps), 0.0f, 0.0f);
        }
    }
    if constexpr (!LDS) {
        if (!w.ok)
            return;
        for_each_run<TOF>(p, g, w,
                          [&](int tile, int sx, int sy, int sz, int len, double s0, double s1,
                              double s2, double slast) {
                              const unsigned long long slot = atomicAdd(tiles + c0 + tile, 1ull);
                              if (!FILL || slot >= nruns)
                                  return; // (never past nruns: the passes find the same runs)
                              put_run(runs + kRunF4 * (int64_t)slot, el, sx, sy, sz, len, s0,
                                      s1, s2, slast);
                          });
    } else {
        for (int i = threadIdx.x; i < ncnt; i += blockDim.x)
            lcnt[i] = 0u;
        __syncthreads();
        auto count = [&](int tile, int, int, int, int, double, double, double, double) {
            atomicAdd(lcnt + c0 + tile, 1u);
        };
        if (w.ok)
            for_each_run<TOF>(p, g, w, count);
        __syncthreads();
        if (!FILL) {
            for (int i = threadIdx.x; i < ncnt; i += blockDim.x)
                if (lcnt[i])
                    atomicAdd(tiles + i, (unsigned long long)lcnt[i]);
            return;
        }
        auto *lbase = reinterpret_cast<unsigned long long *>(lcnt + ((ncnt + 1) & ~1));
        for (int i = threadIdx.x; i < ncnt; i += blockDim.x) {
            const unsigned c = lcnt[i];
            lbase[i] = c ? atomicAdd(tiles + i, (unsigned long long)c) : 0ull;
            lcnt[i] = 0u;
        }
        __syncthreads();
        if (w.ok)
            for_each_run<TOF>(p, g, w,
                              [&](int tile, int sx, int sy, int sz, int len, double s0,
                                  double s1, double s2, double slast) {
                                  const unsigned long long slot =
                                      lbase[c0 + tile] + atomicAdd(lcnt + c0 + tile, 1u);
                                  if (slot >= nruns)
                                      return;
                                  put_run(runs + kRunF4 * (int64_t)slot, el, sx, sy, sz, len,
                                          s0, s1, s2, slast);
                              });
    }
}

// One workgroup per chunk of one tile's runs: zero the FP64 LDS tile, walk
// every run of the chunk into it (scale from vals: 1 / v for v > 0 as
// BackRayTracing's callers, siddon.cu.cc:363-371, or c / v with per-LOR
// counts c; ADJ: g itself, runs of g = 0 skipped), then add the tile's
// non-zero voxels to the image with one atomic each, consecutive lanes on
// consecutive voxels of the image's fastest axis.  vals (vperm) and counts
// (cperm) are read in the plan's LOR order or, through a permutation, in the
// caller's.
template <bool TOF, bool ADJ>
__global__ __launch_bounds__(256) void sid_tile_scatter(SidParams p, SidTileGeom g,
                                                        const SidChunk *__restrict__ chunks,
                                                        const float4 *__restrict__ runs,
                                                        const float4 *__restrict__ rays,
                                                        const float *__restrict__ vals,
                                                        const int *__restrict__ vperm,
                                                        const float *__restrict__ counts,
                                                        const int *__restrict__ cperm)
{
    extern __shared__ double acc[];
    const SidChunk ch = chunks[blockIdx.x];
    for (int v = threadIdx.x; v < g.tvox; v += blockDim.x)
        acc[v] = 0.0;
    const int kz = ch.tile % g.nt[2];
    const int ky = (ch.tile / g.nt[2]) % g.nt[1];
    const int kx = ch.tile / (g.nt[2] * g.nt[1]);
    const int x0 = kx * g.t[0], y0 = ky * g.t[1], z0 = kz * g.t[2];
    __syncthreads();
    for (int j = threadIdx.x; j < ch.count; j += blockDim.x) {
        const float4 *L = runs + kRunF4 * (ch.start + j);
        const float4 a = L[0], b = L[1], c = L[2];
        const int e = __float_as_int(a.x);
        const float v = vals[vperm ? vperm[e] : e];
        float scale;
        if (ADJ) {
            if (v == 0.0f)
                continue;
            scale = v;
        } else {
            if (!(v > 0.0f))
                continue;
            scale = counts ? counts[cperm ? cperm[e] : e] / v : 1.0f / v;
        }
        const float4 *R = rays + kRayF4 * (int64_t)e;
        double t0, t1, t2, t_tof;
        unpack_d2(R[0], t0, t1);
        unpack_d2(R[1], t2, t_tof);
        const float4 rw = R[2];
        const int steps = __float_as_int(rw.y);
        const int pxy = __float_as_int(a.y), pzl = __float_as_int(a.z);
        const int len = (int)((unsigned)pzl >> 16);
        int l = ((pxy & 0xffff) - x0) * g.ls[0] + ((int)((unsigned)pxy >> 16) - y0) * g.ls[1] +
                ((pzl & 0xffff) - z0) * g.ls[2];
        const int d0 = ((steps & 3) - 1) * g.ls[0];
        const int d1 = (((steps >> 2) & 3) - 1) * g.ls[1];
        const int d2 = (((steps >> 4) & 3) - 1) * g.ls[2];
        double n0, n1, n2, last_t;
        unpack_d2(b, n0, n1);
        unpack_d2(c, n2, last_t);
        for (int k = 0; k < len; k++) {
            int i = (n0 <= n1) ? ((n0 <= n2) ? 0 : 2) : ((n1 <= n2) ? 1 : 2);
            double nt = i == 0 ? n0 : (i == 1 ? n1 : n2);
            float gv;
            if (seg_value<TOF>(p, rw.x, t_tof, nt, last_t, gv))
                atomicAdd(acc + l, (double)(gv * scale));
            last_t = nt;
            if (i == 0) {
                n0 = __dadd_rn(nt, t0);
                l += d0;
            } else if (i == 1) {
                n1 = __dadd_rn(nt, t1);
                l += d1;
            } else {
                n2 = __dadd_rn(nt, t2);
                l += d2;
            }
        }
    }
    __syncthreads();
    const int f = g.ord[0], m = g.ord[1], s = g.ord[2];
    const int tf = g.t[f], tm = g.t[m];
    int org[3] = {x0, y0, z0};
    for (int v = threadIdx.x; v < g.tvox; v += blockDim.x) {
        const double sum = acc[v];
        if (sum == 0.0)
            continue;
        int q[3];
        q[f] = v % tf;
        q[m] = (v / tf) % tm;
        q[s] = v / (tf * tm);
        const int ix = org[0] + q[0], iy = org[1] + q[1], iz = org[2] + q[2];
        // a non-zero voxel lies inside the image (the runs stay inside it)
        if (ix < p.g[0] && iy < p.g[1] && iz < p.g[2])
            unsafeAtomicAdd(p.out + ix * p.o[0] + iy * p.o[1] + iz * p.o[2], (float)sum);
    }
}

// Sort key of a LOR for the sorted LOR stream: transaxial direction (8
// bits), tilt (6 bits), signed transaxial distance from the image axis (8
// bits), axial midpoint (10 bits), so that consecutive LORs run nearly
// parallel through neighbouring voxels.  The plan index rides above them, so
// each plan's LORs stay in its own range.  Any order gives the same results
// up to summation order; the key only buys locality.
__global__ __launch_bounds__(256) void sid_sort_key(SidParams p, const int64_t *__restrict__ bounds,
                                                    int nplans, unsigned long long *keys,
                                                    int *idx)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t e = bounds[0] + j;
    if (e >= bounds[nplans])
        return;
    int lo = 0, hi = nplans - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bounds[mid] <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    const float *L = p.lors;
    const int64_t ld = p.ld;
    const float x1 = L[e], y1 = L[ld + e], z1 = L[2 * ld + e];
    float dx = L[3 * ld + e] - x1, dy = L[4 * ld + e] - y1, dz = L[5 * ld + e] - z1;
    if (dx < 0.0f || (dx == 0.0f && dy < 0.0f)) {
        dx = -dx;
        dy = -dy;
        dz = -dz;
    }
    const float lxy = sqrtf(dx * dx + dy * dy);
    const float pi = 3.14159265f;
    const float fphi = lxy > 0.0f ? atan2f(dy, dx) / pi + 0.5f : 0.5f;
    const float fth = lxy > 0.0f ? atanf(dz / lxy) / pi + 0.5f : 1.0f;
    const float rad = 0.5f * sqrtf(p.size[0] * p.size[0] + p.size[1] * p.size[1]);
    const float s = lxy > 0.0f ? ((x1 - p.center[0]) * dy - (y1 - p.center[1]) * dx) / lxy : 0.0f;
    const float fs = s / (2.0f * rad) + 0.5f;
    const float fz = (0.5f * (z1 + L[5 * ld + e]) - p.center[2]) / p.size[2] + 0.5f;
    auto q = [](float f, int bits) {
        const float m = (float)(1 << bits);
        return (unsigned)fminf(fmaxf(f * m, 0.0f), m - 1.0f);
    };
    const unsigned key = (q(fphi, 8) << 24) | (q(fth, 6) << 18) | (q(fs, 8) << 10) | q(fz, 10);
    keys[j] = ((unsigned long long)lo << 32) | key;
    idx[j] = (int)j;
}

// The sorted LOR copy (7 columns, leading dimension n, LOR bounds[0] + j in
// sorted position j) and, for each sorted LOR, its index inside its plan in
// the caller's order.
__global__ __launch_bounds__(256) void sid_sort_gather(SidParams p, const int64_t *__restrict__ bounds,
                                                       int nplans,
                                                       const unsigned long long *__restrict__ keys,
                                                       const int *__restrict__ idx, float *slors,
                                                       int *perm)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t e = bounds[0] + j;
    if (e >= bounds[nplans])
        return;
    const int64_t src = bounds[0] + idx[j];
    const int k = (int)(keys[j] >> 32);
#pragma unroll
    for (int c = 0; c < 7; c++)
        slors[c * p.n + e] = p.lors[c * p.ld + src];
    perm[e] = (int)(src - bounds[k]);
}

// out[perm[e]] = in[e]: a plan's buffer back in the caller's LOR order.
__global__ __launch_bounds__(256) void sid_unsort(const float *__restrict__ in,
                                                  const int *__restrict__ perm, int64_t n,
                                                  float *out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n)
        out[perm[e]] = in[e];
}

} // namespace
} // namespace srfx

using namespace srfx;

struct srfx_siddon_plan_s {
    // device blocks shared by the plans of one set, freed with the last plan
    struct Arena {
        void *mem[3];
        std::atomic<int> refs;
    };
    SidParams p{};
    srfx_grid grid{};
    SidTileGeom g{};
    int nchunks = 0, block = 256;
    int64_t nruns = 0;
    float4 *rays = nullptr;
    float4 *runs = nullptr;
    SidChunk *chunks = nullptr;
    float *proj = nullptr;
    int *perm = nullptr; // sorted plans: plan order -> caller order
    Arena *arena = nullptr;
};

namespace {

void plan_release(srfx_siddon_plan_s *pl)
{
    if (pl->arena && pl->arena->refs.fetch_sub(1) == 1) {
        for (void *m : pl->arena->mem)
            (void)hipFree(m);
        delete pl->arena;
    }
    delete pl;
}

int sid_setup(SidParams &p, const float *lors, int64_t n, int64_t ld, const srfx_grid &g,
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

// LDS bytes per (plan, tile) counter of the LDS plan passes, and the most
// counters they take (64 KB of LDS).
constexpr int kLdsCounterBytes = 12;
constexpr int kMaxLdsCounters = 65536 / kLdsCounterBytes - 1;

template <bool FILL>
void launch_pass(const SidParams &p, const SidTileGeom &g, unsigned long long *tiles,
                 float4 *rays, float4 *runs, int64_t nruns, const int64_t *bounds, int nplans,
                 int ntiles, hipStream_t st)
{
    const dim3 grid(blocks_for(p.n, 256)), block(256);
    const int ncnt = nplans * ntiles;
    const unsigned long long nr = (unsigned long long)nruns;
    if (ncnt <= kMaxLdsCounters) {
        const unsigned lds = (unsigned)(4 * ((ncnt + 1) & ~1) + 8 * ncnt);
        if (p.use_tof)
            hipLaunchKernelGGL((sid_plan_pass<FILL, true, true>), grid, block, lds, st, p, g,
                               tiles, rays, runs, nr, bounds, nplans, ntiles);
        else
            hipLaunchKernelGGL((sid_plan_pass<FILL, false, true>), grid, block, lds, st, p, g,
                               tiles, rays, runs, nr, bounds, nplans, ntiles);
    } else {
        if (p.use_tof)
            hipLaunchKernelGGL((sid_plan_pass<FILL, true, false>), grid, block, 0, st, p, g,
                               tiles, rays, runs, nr, bounds, nplans, ntiles);
        else
            hipLaunchKernelGGL((sid_plan_pass<FILL, false, false>), grid, block, 0, st, p, g,
                               tiles, rays, runs, nr, bounds, nplans, ntiles);
    }
}

// vals in the caller's LOR order (vperm: the plan's sorted order to it) or,
// with vperm null, in the plan's order; counts (optional) likewise (cperm).
int plan_scatter(const srfx_siddon_plan_s *pl, bool adj, const float *vals, const int *vperm,
                 const float *counts, const int *cperm, float *image, uint32_t flags,
                 hipStream_t st)
{
    int rc = zero_grid(image, pl->grid, flags, st);
    if (rc || pl->nchunks == 0)
        return rc;
    SidParams p = pl->p;
    p.out = image;
    const dim3 grid(pl->nchunks), block(pl->block);
    const unsigned lds = (unsigned)(sizeof(double) * pl->g.tvox);
    if (p.use_tof) {
        if (adj)
            hipLaunchKernelGGL((sid_tile_scatter<true, true>), grid, block, lds, st, p, pl->g,
                               pl->chunks, pl->runs, pl->rays, vals, vperm, counts, cperm);
        else
            hipLaunchKernelGGL((sid_tile_scatter<true, false>), grid, block, lds, st, p, pl->g,
                               pl->chunks, pl->runs, pl->rays, vals, vperm, counts, cperm);
    } else {
        if (adj)
            hipLaunchKernelGGL((sid_tile_scatter<false, true>), grid, block, lds, st, p, pl->g,
                               pl->chunks, pl->runs, pl->rays, vals, vperm, counts, cperm);
        else
            hipLaunchKernelGGL((sid_tile_scatter<false, false>), grid, block, lds, st, p, pl->g,
                               pl->chunks, pl->runs, pl->rays, vals, vperm, counts, cperm);
    }
    return launch_status();
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// The plans of rays bounds[k] .. bounds[k + 1] - 1 of one LOR array, built
// together: one count pass and one fill pass over all the rays, two host
// synchronisations and two device blocks for the whole set.
int build_set(const float *lors, int64_t n, int64_t ld, const srfx_grid &grid, float tof_bin,
              float tof_sigma2, const srfx_siddon_tile &tile, int K, const int64_t *bounds,
              srfx_siddon_plan *plans, hipStream_t st)
{
    if (K < 1 || !plans || !bounds)
        return SRFX_EINVAL;
    for (int k = 0; k < K; k++)
        plans[k] = nullptr;
    if (bounds[0] < 0 || bounds[K] > n)
        return SRFX_EINVAL;
    for (int k = 0; k < K; k++)
        if (bounds[k + 1] < bounds[k] || bounds[k + 1] - bounds[k] > (int64_t)INT32_MAX)
            return SRFX_EINVAL;
    const int t[3] = {tile.tx, tile.ty, tile.tz};
    for (int c = 0; c < 3; c++)
        if (t[c] <= 0 || grid.shape[c] >= 65536)
            return SRFX_EINVAL;
    if ((int64_t)t[0] * t[1] * t[2] > kMaxTileVox || tile.chunk < 0 ||
        !(tile.block == 0 || tile.block == 64 || tile.block == 128 || tile.block == 256))
        return SRFX_EINVAL;
    SidParams p;
    int rc = sid_setup(p, lors, n, ld, grid, tof_bin, tof_sigma2);
    if (rc)
        return rc;
    const int chunk = tile.chunk ? tile.chunk : 2048;
    SidTileGeom g{};
    for (int c = 0; c < 3; c++) {
        g.t[c] = std::min(t[c], grid.shape[c]);
        g.nt[c] = (grid.shape[c] + g.t[c] - 1) / g.t[c];
    }
    // LDS order of a tile's voxels: the image's own axis order (fastest
    // output stride first), so that a flush writes runs of consecutive voxels
    int ord[3] = {0, 1, 2};
    std::stable_sort(ord, ord + 3, [&](int a, int b) { return p.o[a] < p.o[b]; });
    int ls = 1;
    for (int k = 0; k < 3; k++) {
        g.ord[k] = ord[k];
        g.ls[ord[k]] = ls;
        ls *= g.t[ord[k]];
    }
    g.tvox = ls;
    const int ntiles = g.nt[0] * g.nt[1] * g.nt[2];
    const int64_t ncnt = (int64_t)K * ntiles;
    // block 0: ray records, projection buffer, (plan, tile) counters, bounds
    const size_t nr = (size_t)std::max<int64_t>(n, 1);
    const size_t o_proj = align256(sizeof(float4) * kRayF4 * nr);
    const size_t o_cnt = o_proj + align256(sizeof(float) * nr);
    const size_t o_bnd = o_cnt + align256(sizeof(unsigned long long) * ncnt);
    const size_t b0 = o_bnd + align256(sizeof(int64_t) * (K + 1));
    char *blk0 = nullptr, *blk1 = nullptr, *blk2 = nullptr;
    if (hipMalloc(&blk0, b0) != hipSuccess)
        return SRFX_ENOMEM;
    auto fail = [&](int code) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(blk0);
        (void)hipFree(blk1);
        (void)hipFree(blk2);
        return code;
    };
    float4 *rays = (float4 *)blk0;
    float *proj = (float *)(blk0 + o_proj);
    auto *cnt = (unsigned long long *)(blk0 + o_cnt);
    auto *dbounds = (int64_t *)(blk0 + o_bnd);
    std::vector<unsigned long long> hc(ncnt, 0ull);
    const bool any = bounds[K] > bounds[0];
    const int64_t m = bounds[K] - bounds[0];
    int *perm = nullptr;
    if (any && (hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * ncnt, st) != hipSuccess ||
                hipMemcpyAsync(dbounds, bounds, sizeof(int64_t) * (K + 1),
                               hipMemcpyHostToDevice, st) != hipSuccess))
        return fail(SRFX_ELAUNCH);
    if (any && tile.sort) {
        // the sorted LOR stream: keys, a radix sort of (key, index), then the
        // sorted copy of the LORs that every pass and projection reads
        if (m > (int64_t)INT32_MAX - 1)
            return fail(SRFX_EINVAL);
        size_t tmp_bytes = 0;
        const int key_bits = 32 + std::max(1, 32 - __builtin_clz((unsigned)K));
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, (unsigned long long *)nullptr,
                                               (unsigned long long *)nullptr, (int *)nullptr,
                                               (int *)nullptr, (int)m, 0, key_bits, st) !=
            hipSuccess)
            return fail(SRFX_ELAUNCH);
        const size_t o_perm = align256(sizeof(float) * 7 * nr);
        const size_t o_k0 = o_perm + align256(sizeof(int) * nr);
        const size_t o_k1 = o_k0 + align256(sizeof(unsigned long long) * m);
        const size_t o_i0 = o_k1 + align256(sizeof(unsigned long long) * m);
        const size_t o_i1 = o_i0 + align256(sizeof(int) * m);
        const size_t o_tmp = o_i1 + align256(sizeof(int) * m);
        if (hipMalloc(&blk2, o_tmp + tmp_bytes) != hipSuccess)
            return fail(SRFX_ENOMEM);
        auto *slors = (float *)blk2;
        perm = (int *)(blk2 + o_perm);
        auto *k0 = (unsigned long long *)(blk2 + o_k0), *k1 = (unsigned long long *)(blk2 + o_k1);
        auto *i0 = (int *)(blk2 + o_i0), *i1 = (int *)(blk2 + o_i1);
        const dim3 sg(blocks_for(m, 256)), sb(256);
        hipLaunchKernelGGL(sid_sort_key, sg, sb, 0, st, p, dbounds, K, k0, i0);
        if (launch_status() ||
            hipcub::DeviceRadixSort::SortPairs(blk2 + o_tmp, tmp_bytes, k0, k1, i0, i1, (int)m, 0,
                                               key_bits, st) != hipSuccess)
            return fail(SRFX_ELAUNCH);
        hipLaunchKernelGGL(sid_sort_gather, sg, sb, 0, st, p, dbounds, K, k1, i1, slors, perm);
        if (launch_status())
            return fail(SRFX_ELAUNCH);
        p.lors = slors;
        p.ld = n;
    }
    if (any) {
        launch_pass<false>(p, g, cnt, rays, nullptr, 0, dbounds, K, ntiles, st);
        if (launch_status() ||
            hipMemcpyAsync(hc.data(), cnt, sizeof(unsigned long long) * ncnt,
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return fail(SRFX_ELAUNCH);
    }
    // (plan, tile) offsets into one run array, then each plan's chunks of at
    // most `chunk` runs of one tile (starts relative to the plan's runs)
    std::vector<unsigned long long> off(ncnt);
    std::vector<SidChunk> chunks;
    std::vector<int64_t> run0(K + 1), chunk0(K + 1);
    unsigned long long total = 0;
    for (int k = 0; k < K; k++) {
        run0[k] = (int64_t)total;
        chunk0[k] = (int64_t)chunks.size();
        for (int tt = 0; tt < ntiles; tt++) {
            const unsigned long long c = hc[(size_t)k * ntiles + tt];
            off[(size_t)k * ntiles + tt] = total;
            for (unsigned long long s0 = 0; s0 < c; s0 += (unsigned long long)chunk)
                chunks.push_back({(int64_t)(total + s0) - run0[k], tt,
                                  (int)std::min<unsigned long long>(chunk, c - s0)});
            total += c;
        }
    }
    run0[K] = (int64_t)total;
    chunk0[K] = (int64_t)chunks.size();
    float4 *runs = nullptr;
    SidChunk *dchunks = nullptr;
    if (total > 0) {
        const size_t o_ch = align256(sizeof(float4) * kRunF4 * total);
        if (hipMalloc(&blk1, o_ch + sizeof(SidChunk) * chunks.size()) != hipSuccess)
            return fail(SRFX_ENOMEM);
        runs = (float4 *)blk1;
        dchunks = (SidChunk *)(blk1 + o_ch);
        if (hipMemcpyAsync(cnt, off.data(), sizeof(unsigned long long) * ncnt,
                           hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(dchunks, chunks.data(), sizeof(SidChunk) * chunks.size(),
                           hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(SRFX_ELAUNCH);
        launch_pass<true>(p, g, cnt, rays, runs, (int64_t)total, dbounds, K, ntiles, st);
        if (launch_status())
            return fail(SRFX_ELAUNCH);
    }
    if (hipStreamSynchronize(st) != hipSuccess)
        return fail(SRFX_ELAUNCH);
    auto *arena = new srfx_siddon_plan_s::Arena;
    arena->mem[0] = blk0;
    arena->mem[1] = blk1;
    arena->mem[2] = blk2;
    arena->refs.store(K);
    for (int k = 0; k < K; k++) {
        auto *pl = new srfx_siddon_plan_s;
        pl->p = p;
        pl->p.lors = n > 0 ? p.lors + bounds[k] : p.lors;
        pl->perm = perm ? perm + bounds[k] : nullptr;
        pl->p.n = bounds[k + 1] - bounds[k];
        pl->grid = grid;
        pl->g = g;
        pl->block = tile.block ? tile.block : 256;
        pl->rays = rays + kRayF4 * bounds[k];
        pl->proj = proj + bounds[k];
        pl->runs = runs ? runs + kRunF4 * run0[k] : nullptr;
        pl->nruns = run0[k + 1] - run0[k];
        pl->chunks = dchunks ? dchunks + chunk0[k] : nullptr;
        pl->nchunks = (int)(chunk0[k + 1] - chunk0[k]);
        pl->arena = arena;
        plans[k] = pl;
    }
    return SRFX_OK;
}

} // namespace

extern "C" int srfx_siddon_plan_create(const float *lors, int64_t n, int64_t ld, srfx_grid grid,
                                       float tof_bin, float tof_sigma2, srfx_siddon_tile tile,
                                       srfx_siddon_plan *plan, void *stream)
{
    if (!plan)
        return SRFX_EINVAL;
    const int64_t bounds[2] = {0, n < 0 ? 0 : n};
    if (n < 0)
        return SRFX_EINVAL;
    return build_set(lors, n, ld, grid, tof_bin, tof_sigma2, tile, 1, bounds, plan,
                     (hipStream_t)stream);
}

extern "C" int srfx_siddon_plan_create_subsets(const float *lors, int64_t n, int64_t ld,
                                               srfx_grid grid, float tof_bin, float tof_sigma2,
                                               srfx_siddon_tile tile, int32_t nplans,
                                               const int64_t *bounds, srfx_siddon_plan *plans,
                                               void *stream)
{
    return build_set(lors, n, ld, grid, tof_bin, tof_sigma2, tile, nplans, bounds, plans,
                     (hipStream_t)stream);
}

extern "C" int srfx_siddon_plan_info(srfx_siddon_plan plan, int64_t *nruns, int32_t *nchunks)
{
    if (!plan)
        return SRFX_EINVAL;
    if (nruns)
        *nruns = plan->nruns;
    if (nchunks)
        *nchunks = plan->nchunks;
    return SRFX_OK;
}

extern "C" int srfx_siddon_plan_destroy(srfx_siddon_plan plan)
{
    if (!plan)
        return SRFX_OK;
    (void)hipDeviceSynchronize();
    plan_release(plan);
    return SRFX_OK;
}

extern "C" int srfx_siddon_plan_backproject(srfx_siddon_plan plan, const float *lors_value,
                                            float *image, uint32_t flags, void *stream)
{
    if (!plan || !image || (plan->p.n > 0 && !lors_value))
        return SRFX_EINVAL;
    return plan_scatter(plan, false, lors_value, plan->perm, nullptr, nullptr, image, flags,
                        (hipStream_t)stream);
}

extern "C" int srfx_siddon_plan_backproject_adjoint(srfx_siddon_plan plan, const float *g,
                                                    float *image, uint32_t flags, void *stream)
{
    if (!plan || !image || (plan->p.n > 0 && !g))
        return SRFX_EINVAL;
    return plan_scatter(plan, true, g, plan->perm, nullptr, nullptr, image, flags,
                        (hipStream_t)stream);
}

extern "C" int srfx_siddon_plan_osem_subset(srfx_siddon_plan plan, const float *image,
                                            const float *counts, float *bp, uint32_t flags,
                                            void *stream)
{
    const bool project = !(flags & SRFX_OSEM_SCATTER_ONLY);
    const bool scatter = !(flags & SRFX_OSEM_PROJECT_ONLY);
    if (!plan || (scatter && !bp) || (project && plan->p.n > 0 && !image) ||
        (!project && !scatter))
        return SRFX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const SidParams &p = plan->p;
    if (project && p.n > 0) {
        const int rc = srfx_siddon_project(p.lors, p.n, p.ld, image, plan->grid, p.tof_bin,
                                           p.tof_sigma2, plan->proj, 0, stream);
        if (rc)
            return rc;
    }
    if (!scatter)
        return SRFX_OK;
    return plan_scatter(plan, false, plan->proj, nullptr, counts, plan->perm, bp,
                        flags & SRFX_ACCUMULATE, st);
}

extern "C" int srfx_siddon_plan_projection(srfx_siddon_plan plan, float *out, void *stream)
{
    if (!plan || (plan->p.n > 0 && !out))
        return SRFX_EINVAL;
    if (plan->p.n == 0)
        return SRFX_OK;
    hipStream_t st = (hipStream_t)stream;
    if (plan->perm) {
        hipLaunchKernelGGL(sid_unsort, dim3(blocks_for(plan->p.n, 256)), dim3(256), 0, st,
                           (const float *)plan->proj, (const int *)plan->perm, plan->p.n, out);
        return launch_status();
    }
    if (hipMemcpyAsync(out, plan->proj, sizeof(float) * plan->p.n, hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
        return SRFX_ELAUNCH;
    return SRFX_OK;
}
