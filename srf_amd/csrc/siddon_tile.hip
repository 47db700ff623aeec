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

// The plan set's two passes over the rays of all its plans (plan k holds
// rays bounds[k] .. bounds[k + 1] - 1), one lane per ray: the walk of
// siddon_core.h's walk() (same argmin, same FP64 updates, same end
// conditions), cut into runs at tile boundaries.  A run with at least one
// kept segment (all of them without TOF; t^2 / sigma^2 < 9 with TOF) is a
// listing.  Count pass (FILL false): writes the ray records and counts the
// listings of each (plan, tile).  Fill pass: writes each listing at its
// (plan, tile) cursor, with the ray's index inside its plan.
template <bool FILL, bool TOF>
__global__ __launch_bounds__(256) void sid_plan_pass(SidParams p, SidTileGeom g,
                                                     unsigned long long *tiles, float4 *rays,
                                                     float4 *runs, unsigned long long nruns,
                                                     const int64_t *__restrict__ bounds,
                                                     int nplans, int ntiles)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p.n || e < bounds[0] || e >= bounds[nplans])
        return;
    // the plan of ray e: the last k with bounds[k] <= e
    int lo = 0, hi = nplans - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bounds[mid] <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    unsigned long long *ptiles = tiles + (int64_t)lo * ntiles;
    const int el = (int)(e - bounds[lo]);
    Walk w;
    setup_walk<kBack>(p, e, w);
    if (!FILL) {
        float4 *r = rays + kRayF4 * e;
        r[0] = pack_d2(w.ok ? w.deltaT[0] : 0.0, w.ok ? w.deltaT[1] : 0.0);
        r[1] = pack_d2(w.ok ? w.deltaT[2] : 0.0, w.ok ? w.t_tof : 0.0);
        const int steps = w.ok ? (w.step[0] + 1) | ((w.step[1] + 1) << 2) | ((w.step[2] + 1) << 4)
                               : 0;
        r[2] = make_float4(w.weight, __int_as_float(steps), 0.0f, 0.0f);
    }
    if (!w.ok)
        return;
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
    auto emit = [&]() {
        const int tile = tile_id(g, sx, sy, sz);
        if (!FILL) {
            atomicAdd(ptiles + tile, 1ull);
        } else {
            const unsigned long long slot = atomicAdd(ptiles + tile, 1ull);
            if (slot >= nruns)
                return; // never: the fill pass finds the runs the count pass counted
            float4 *o = runs + kRunF4 * (int64_t)slot;
            o[0] = make_float4(__int_as_float(el), __int_as_float(sx | (sy << 16)),
                               __int_as_float(sz | (len << 16)), 0.0f);
            o[1] = pack_d2(s0, s1);
            o[2] = pack_d2(s2, slast);
        }
    };
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
                emit();
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
        emit();
}

// One workgroup per chunk of one tile's runs: zero the FP64 LDS tile, walk
// every run of the chunk into it (scale from vals: 1 / v for v > 0 as
// BackRayTracing's callers, siddon.cu.cc:363-371; ADJ: g itself, runs of
// g = 0 skipped), then add the tile's non-zero voxels to the image with one
// atomic each, consecutive lanes on consecutive voxels of the image's
// fastest axis.
template <bool TOF, bool ADJ>
__global__ __launch_bounds__(256) void sid_tile_scatter(SidParams p, SidTileGeom g,
                                                        const SidChunk *__restrict__ chunks,
                                                        const float4 *__restrict__ runs,
                                                        const float4 *__restrict__ rays,
                                                        const float *__restrict__ vals)
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
        const float v = vals[e];
        float scale;
        if (ADJ) {
            if (v == 0.0f)
                continue;
            scale = v;
        } else {
            if (!(v > 0.0f))
                continue;
            scale = 1.0f / v;
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

} // namespace
} // namespace srfx

using namespace srfx;

struct srfx_siddon_plan_s {
    // device blocks shared by the plans of one set, freed with the last plan
    struct Arena {
        void *mem[2];
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
    Arena *arena = nullptr;
};

namespace {

void plan_release(srfx_siddon_plan_s *pl)
{
    if (pl->arena && pl->arena->refs.fetch_sub(1) == 1) {
        (void)hipFree(pl->arena->mem[0]);
        (void)hipFree(pl->arena->mem[1]);
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

template <bool FILL>
void launch_pass(const SidParams &p, const SidTileGeom &g, unsigned long long *tiles,
                 float4 *rays, float4 *runs, int64_t nruns, const int64_t *bounds, int nplans,
                 int ntiles, hipStream_t st)
{
    const dim3 grid(blocks_for(p.n, 256)), block(256);
    if (p.use_tof)
        hipLaunchKernelGGL((sid_plan_pass<FILL, true>), grid, block, 0, st, p, g, tiles, rays,
                           runs, (unsigned long long)nruns, bounds, nplans, ntiles);
    else
        hipLaunchKernelGGL((sid_plan_pass<FILL, false>), grid, block, 0, st, p, g, tiles, rays,
                           runs, (unsigned long long)nruns, bounds, nplans, ntiles);
}

int plan_scatter(const srfx_siddon_plan_s *pl, bool adj, const float *vals, float *image,
                 uint32_t flags, hipStream_t st)
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
                               pl->chunks, pl->runs, pl->rays, vals);
        else
            hipLaunchKernelGGL((sid_tile_scatter<true, false>), grid, block, lds, st, p, pl->g,
                               pl->chunks, pl->runs, pl->rays, vals);
    } else {
        if (adj)
            hipLaunchKernelGGL((sid_tile_scatter<false, true>), grid, block, lds, st, p, pl->g,
                               pl->chunks, pl->runs, pl->rays, vals);
        else
            hipLaunchKernelGGL((sid_tile_scatter<false, false>), grid, block, lds, st, p, pl->g,
                               pl->chunks, pl->runs, pl->rays, vals);
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
    char *blk0 = nullptr, *blk1 = nullptr;
    if (hipMalloc(&blk0, b0) != hipSuccess)
        return SRFX_ENOMEM;
    auto fail = [&](int code) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(blk0);
        (void)hipFree(blk1);
        return code;
    };
    float4 *rays = (float4 *)blk0;
    float *proj = (float *)(blk0 + o_proj);
    auto *cnt = (unsigned long long *)(blk0 + o_cnt);
    auto *dbounds = (int64_t *)(blk0 + o_bnd);
    std::vector<unsigned long long> hc(ncnt, 0ull);
    const bool any = bounds[K] > bounds[0];
    if (any) {
        if (hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * ncnt, st) != hipSuccess ||
            hipMemcpyAsync(dbounds, bounds, sizeof(int64_t) * (K + 1), hipMemcpyHostToDevice,
                           st) != hipSuccess)
            return fail(SRFX_ELAUNCH);
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
    arena->refs.store(K);
    for (int k = 0; k < K; k++) {
        auto *pl = new srfx_siddon_plan_s;
        pl->p = p;
        pl->p.lors = n > 0 ? lors + bounds[k] : lors;
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
    return plan_scatter(plan, false, lors_value, image, flags, (hipStream_t)stream);
}

extern "C" int srfx_siddon_plan_backproject_adjoint(srfx_siddon_plan plan, const float *g,
                                                    float *image, uint32_t flags, void *stream)
{
    if (!plan || !image || (plan->p.n > 0 && !g))
        return SRFX_EINVAL;
    return plan_scatter(plan, true, g, image, flags, (hipStream_t)stream);
}

extern "C" int srfx_siddon_plan_osem_subset(srfx_siddon_plan plan, const float *image,
                                            float *bp, uint32_t flags, void *stream)
{
    if (!plan || !bp || (plan->p.n > 0 && !image))
        return SRFX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    const SidParams &p = plan->p;
    if (p.n > 0) {
        const int rc = srfx_siddon_project(p.lors, p.n, p.ld, image, plan->grid, p.tof_bin,
                                           p.tof_sigma2, plan->proj, 0, stream);
        if (rc)
            return rc;
    }
    return plan_scatter(plan, false, plan->proj, bp, flags, st);
}

extern "C" int srfx_siddon_plan_projection(srfx_siddon_plan plan, float *out, void *stream)
{
    if (!plan || (plan->p.n > 0 && !out))
        return SRFX_EINVAL;
    if (plan->p.n > 0 &&
        hipMemcpyAsync(out, plan->proj, sizeof(float) * plan->p.n, hipMemcpyDeviceToDevice,
                       (hipStream_t)stream) != hipSuccess)
        return SRFX_ELAUNCH;
    return SRFX_OK;
}
