// Tiled TOR ops: a per-event-set plan that bins every event into the image
// tiles it can touch, then gathers from / scatters into tiles staged in LDS.
//
// A tile is a slab of S consecutive slices times a window of Wx x Wy voxels
// of the op frame.  Along a TOR group's dominant axis an event moves by at
// most one voxel pitch per slice, so within one slab its voxels stay inside a
// small footprint and it touches few tiles.  The plan lists, for every tile,
// the events whose footprint meets it ("listings"), cut into chunks of at most
// `chunk` listings; one workgroup takes one chunk:
//   scatter: zero the tile (FP64) in LDS, every lane walks one listing's
//            voxels that lie in the tile and adds its weights with LDS
//            atomics, then the tile is flushed with one coalesced FP32 global
//            atomic per non-zero voxel (instead of one scattered global atomic
//            per weight);
//   gather:  load the tile into LDS, every lane sums its listing's weighted
//            voxels from LDS into one partial per listing; a second pass sums
//            each event's partials in the event's own listing order, so the
//            projection is deterministic.
// The voxels an event visits and their weights are those of the untiled
// kernels (tor_core.h): the tile walk is the candidate-box (or chord) walk
// clipped to the window.
#include <rocprim/device/device_scan.hpp>

#include <vector>

#include "tor_core.h"

namespace srfx {
namespace {

struct TileGeom {
    int S, Wx, Wy;       // slices per slab, window width along op x and y (voxels)
    int nwx, nwy, nslab; // windows per axis and number of slabs
};

struct TileChunk {
    int tile, start, count, pad;
};

__device__ __forceinline__ int trunc_clamped(double u)
{
    return (int)fmin(fmax(u, -1.0e9), 1.0e9);
}

// Windows wx0..wx1, wy0..wy1 holding every voxel the event can visit in the
// slices s0..s1.  The patch origin of slice s is (int)u(s) - half with u
// affine in s (cross_at), or the reference's exact origin, which differs from
// it by at most one voxel; the patch spans P voxels from its origin.
__device__ __forceinline__ bool footprint(const TorParams &p, const TileGeom &g, const Track &t,
                                          int s0, int s1, int &wx0, int &wx1, int &wy0, int &wy1)
{
    int a0 = trunc_clamped(t.ax + (double)s0 * t.bx) - p.half;
    int a1 = trunc_clamped(t.ax + (double)s1 * t.bx) - p.half;
    int b0 = trunc_clamped(t.ay + (double)s0 * t.by) - p.half;
    int b1 = trunc_clamped(t.ay + (double)s1 * t.by) - p.half;
    int x0 = max(min(a0, a1) - 1, 0), x1 = min(max(a0, a1) + p.patch, p.gx - 1);
    int y0 = max(min(b0, b1) - 1, 0), y1 = min(max(b0, b1) + p.patch, p.gy - 1);
    if (x0 > x1 || y0 > y1)
        return false;
    wx0 = x0 / g.Wx;
    wx1 = x1 / g.Wx;
    wy0 = y0 / g.Wy;
    wy1 = y1 / g.Wy;
    return true;
}

// Calls f(tile) for every tile the event can touch: slabs in slice order,
// then windows row by row.  The count and fill passes enumerate identically.
template <class F>
__device__ __forceinline__ void for_each_tile(const TorParams &p, const TileGeom &g,
                                              const Event &ev, const Track &t, F f)
{
    const int sl0 = ev.s_lo / g.S, sl1 = ev.s_hi / g.S;
    for (int sl = sl0; sl <= sl1; sl++) {
        int s0 = max(ev.s_lo, sl * g.S), s1 = min(ev.s_hi, sl * g.S + g.S - 1);
        int wx0, wx1, wy0, wy1;
        if (!footprint(p, g, t, s0, s1, wx0, wx1, wy0, wy1))
            continue;
        for (int wy = wy0; wy <= wy1; wy++)
            for (int wx = wx0; wx <= wx1; wx++)
                f((sl * g.nwy + wy) * g.nwx + wx);
    }
}

struct PlanArgs {
    int cols7, ntiles;
    int *tile_cnt;        // count pass: listings per tile; fill pass: read
    int *ev_cnt;          // count pass: listings per event (nullable)
    float4 *rec;          // count pass writes, fill pass reads: 48-byte event records
    const int *ev_off;    // fill pass: first listing of each event (nullable)
    const int *tile_off;  // fill pass: first slot of each tile
    int *tile_fill;       // fill pass: slots taken per tile
    int *list;            // fill pass: event of each slot
    int *ev_slot;         // fill pass: slot of each of an event's listings (nullable)
    int *err;             // fill pass: count of listings that did not fit
};

__device__ __forceinline__ bool plan_init(const TorParams &p, int cols7, const RawEvent &r,
                                          Event &ev)
{
    return cols7 ? init_event<kMap>(p, r, ev) : init_event<kProj>(p, r, ev);
}

// Count pass (FILL false): pack the event's record, count its tiles per
// event and per tile.  Fill pass (FILL true): list the event in each of its
// tiles.  With LDS, a workgroup first counts its own listings per tile in
// LDS, then adds them to the global counters (count pass) or reserves a
// contiguous range of slots per tile with one global atomic (fill pass), so
// the global atomics are per (workgroup, tile) instead of per listing and a
// workgroup's slots of one tile are contiguous.  Both passes enumerate the
// tiles with the same code; the fill pass checks every slot against the
// counted ranges and reports a mismatch in *err instead of writing out of
// range.
template <bool FILL, bool LDS>
__global__ __launch_bounds__(256) void plan_pass(TorParams p, TileGeom g, PlanArgs a)
{
    extern __shared__ int sh[]; // LDS: counts [ntiles], FILL: + bases, ranks [ntiles] each
    const int nt = a.ntiles;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = e < p.n;
    if (LDS) {
        for (int i = threadIdx.x; i < (FILL ? 3 : 1) * nt; i += blockDim.x)
            sh[i] = 0;
        __syncthreads();
    }
    RawEvent r;
    Event ev;
    Track t;
    bool ok = false;
    if (valid) {
        if (FILL) {
            fetch_event_rec(a.rec, e, r);
        } else {
            if (a.cols7)
                fetch_event<kMap>(p, e, r);
            else
                fetch_event<kProj>(p, e, r);
            a.rec[3 * e] = make_float4(r.x1, r.y1, r.z1, r.x2);
            a.rec[3 * e + 1] = make_float4(r.y2, r.z2, r.xc, r.yc);
            a.rec[3 * e + 2] = make_float4(r.zc, r.f, 0.0f, 0.0f);
        }
        ok = plan_init(p, a.cols7, r, ev);
        if (ok)
            make_track(p, ev, t);
    }
    if (!FILL) {
        int k = 0;
        if (ok)
            for_each_tile(p, g, ev, t, [&](int tile) {
                atomicAdd(LDS ? &sh[tile] : &a.tile_cnt[tile], 1);
                k++;
            });
        if (valid && a.ev_cnt)
            a.ev_cnt[e] = k;
        if (LDS) {
            __syncthreads();
            for (int i = threadIdx.x; i < nt; i += blockDim.x) {
                int c = sh[i];
                if (c)
                    atomicAdd(&a.tile_cnt[i], c);
            }
        }
        return;
    }
    if (LDS) {
        if (ok)
            for_each_tile(p, g, ev, t, [&](int tile) { atomicAdd(&sh[tile], 1); });
        __syncthreads();
        for (int i = threadIdx.x; i < nt; i += blockDim.x) {
            int c = sh[i];
            sh[nt + i] = c ? atomicAdd(&a.tile_fill[i], c) : 0;
        }
        __syncthreads();
    }
    const int k0 = (valid && a.ev_slot) ? a.ev_off[e] : 0;
    const int k1 = (valid && a.ev_slot) ? a.ev_off[e + 1] : 0;
    int k = 0;
    if (ok)
        for_each_tile(p, g, ev, t, [&](int tile) {
            int rank;
            bool fits;
            if (LDS) {
                int loc = atomicAdd(&sh[2 * nt + tile], 1);
                rank = sh[nt + tile] + loc;
                fits = loc < sh[tile];
            } else {
                rank = atomicAdd(&a.tile_fill[tile], 1);
                fits = true;
            }
            fits = fits && rank < a.tile_cnt[tile] && (!a.ev_slot || k0 + k < k1);
            if (fits) {
                int slot = a.tile_off[tile] + rank;
                a.list[slot] = (int)e;
                if (a.ev_slot)
                    a.ev_slot[k0 + k] = slot;
            } else {
                atomicAdd(a.err, 1);
            }
            k++;
        });
    if (valid && a.ev_slot && k0 + k != k1)
        atomicAdd(a.err, 1);
}

__global__ __launch_bounds__(256) void fill_int(int *v, int64_t n, int value)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        v[i] = value;
}

// Tile bounds of a chunk: slices sa..sb, op-frame voxels x0..x1, y0..y1.
struct TileBox {
    int sa, sb, x0, x1, y0, y1;
};

__device__ __forceinline__ TileBox tile_box(const TorParams &p, const TileGeom &g, int tile)
{
    TileBox b;
    const int per = g.nwx * g.nwy;
    const int sl = tile / per, r = tile - sl * per;
    const int wy = r / g.nwx, wx = r - wy * g.nwx;
    b.sa = sl * g.S;
    b.sb = min(b.sa + g.S, p.gz) - 1;
    b.x0 = wx * g.Wx;
    b.x1 = min(b.x0 + g.Wx, p.gx) - 1;
    b.y0 = wy * g.Wy;
    b.y1 = min(b.y0 + g.Wy, p.gy) - 1;
    return b;
}

// The chord walk of one slice (patch_walk) clipped to the tile's window.
//   KIND 0: returns sum w * lds[voxel]; KIND 1: lds[voxel] += w * scale.
template <int KIND, typename T>
__device__ __forceinline__ float window_walk(const TorParams &p, const TileGeom &g,
                                             const TileBox &b, const Event &ev, const Track &t,
                                             const Cross &c, float amp, int s, float scale,
                                             T *lds)
{
    const int P = p.patch;
    int i_lo = max(0, b.x0 - c.ix0), i_hi = min(P - 1, b.x1 - c.ix0);
    int j_lo = max(0, b.y0 - c.iy0), j_hi = min(P - 1, b.y1 - c.iy0);
    const float eps = 1.0e-4f;
    j_lo = max(j_lo, (int)ceilf((c.dy0 - t.dymax) * p.rfinter_y - eps));
    j_hi = min(j_hi, (int)floorf((c.dy0 + t.dymax) * p.rfinter_y + eps));
    const int row0 = ((s - b.sa) * g.Wy + (c.iy0 - b.y0)) * g.Wx + (c.ix0 - b.x0);
    float acc = 0.0f;
    for (int j = j_lo; j <= j_hi; j++) {
        float dy = voxel_offset(c.dy0, j, p.finter_y);
        float tt = t.tm - dy * dy * t.q;
        if (tt <= 0.0f)
            continue;
        float hw = __builtin_amdgcn_sqrtf(tt * t.A);
        float ctr = c.dx0 - t.cxy * dy;
        int il = max(i_lo, (int)ceilf((ctr - hw) * p.rfinter_x - eps));
        int ih = min(i_hi, (int)floorf((ctr + hw) * p.rfinter_x + eps));
        for (int i = il; i <= ih; i++) {
            float dx = voxel_offset(c.dx0, i, p.finter_x);
            float d2 = tube_dist2(dx, dy, ev.ux, ev.uy);
            bool in = d2 < ev.thr - kThrMargin;
            if (!in && d2 < ev.thr + kThrMargin)
                in = inside_exact(p, ev, s, c.ix0 + i, c.iy0 + j);
            if (in) {
                float w = amp * __builtin_amdgcn_exp2f(d2 * ev.cw);
                int li = row0 + j * g.Wx + i;
                if constexpr (KIND == 0)
                    acc = fmaf(w, lds[li], acc);
                else
                    atomicAdd(&lds[li], (T)(w * scale));
            }
        }
    }
    return acc;
}

// One listing: the event's slices inside the tile, each through the NB x NB
// candidate box clipped to the window when the event's box fits, else the
// clipped chord walk.
template <int KIND, int NB, typename T>
__device__ __forceinline__ float tile_event(const TorParams &p, const TileGeom &g,
                                            const TileBox &b, const Event &ev, bool tof,
                                            float scale, T *lds)
{
    Track t;
    make_track(p, ev, t);
    const bool boxed = NB > 0 && t.nx <= NB && t.ny <= NB;
    const int s0 = max(ev.s_lo, b.sa), s1 = min(ev.s_hi, b.sb);
    const int P = p.patch;
    float acc = 0.0f;
    for (int s = s0; s <= s1; s++) {
        Cross c;
        cross_at(p, ev, t, s, c);
        float amp = tof ? tof_factor(p, ev, t, s) : 1.0f;
        if (NB > 0 && boxed) {
            PatchIdx idx;
            idx.base = ((s - b.sa) * g.Wy + (c.iy0 - b.y0)) * g.Wx + (c.ix0 - b.x0);
            idx.si = 1;
            idx.sj = g.Wx;
            const float *src = nullptr;
            if constexpr (KIND == 0)
                src = lds;
            acc += slice_box<KIND, (NB > 0 ? NB : 1), true>(
                p, ev, t, c, amp, s, scale, max(0, b.x0 - c.ix0), min(P - 1, b.x1 - c.ix0),
                max(0, b.y0 - c.iy0), min(P - 1, b.y1 - c.iy0), idx, src, lds);
        } else {
            acc += window_walk<KIND>(p, g, b, ev, t, c, amp, s, scale, lds);
        }
    }
    return acc;
}

// Scatter of one chunk (modes kBack, kMap, kAdj) into p.out.
template <int MODE, int NB>
__global__ __launch_bounds__(256) void tile_scatter(TorParams p, TileGeom g,
                                                    const TileChunk *__restrict__ chunks,
                                                    const int *__restrict__ list,
                                                    const float4 *__restrict__ rec)
{
    // FP64 tile: LDS FP64 atomic adds run at several lanes per cycle where
    // FP32 ones are serialised lane by lane, and the tile sums gain precision.
    extern __shared__ double ldsd[];
    const TileChunk ch = chunks[blockIdx.x];
    const TileBox b = tile_box(p, g, ch.tile);
    const int nv = g.S * g.Wy * g.Wx;
    for (int i = threadIdx.x; i < nv; i += blockDim.x)
        ldsd[i] = 0.0;
    __syncthreads();
    const bool tof = MODE != kMap && p.use_tof;
    // The 64 lanes of a wave take listings `stride` apart in the chunk:
    // neighbours in the list are neighbouring (often near-identical) events,
    // whose LDS atomics would hit the same voxels in the same instruction
    // and serialise.
    const int stride = (ch.count + 63) >> 6;
    for (int q = threadIdx.x; q < 64 * stride; q += blockDim.x) {
        const int j = (q & 63) * stride + (q >> 6);
        if (j >= ch.count)
            continue;
        const int64_t e = list[ch.start + j];
        float scale;
        bool go = true;
        if (MODE == kBack) {
            float v = p.val[e];
            go = v > 1e-7;
            scale = 1.0f / v;
        } else if (MODE == kMap) {
            scale = 1.0f / p.val[e];
        } else {
            scale = p.val[e];
            go = scale != 0.0f;
        }
        if (go) {
            RawEvent r;
            Event ev;
            fetch_event_rec(rec, e, r);
            if (init_event<MODE>(p, r, ev))
                tile_event<1, NB>(p, g, b, ev, tof, scale, ldsd);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nv; i += blockDim.x) {
        float v = (float)ldsd[i];
        if (v != 0.0f) {
            int lx = i % g.Wx, r = i / g.Wx;
            int ly = r % g.Wy, ls = r / g.Wy;
            int ix = b.x0 + lx, iy = b.y0 + ly, s = b.sa + ls;
            if (ix <= b.x1 && iy <= b.y1 && s <= b.sb)
                unsafeAtomicAdd(p.out + (ix * p.ox + iy * p.oy + s * p.oz), v);
        }
    }
}

// Gather of one chunk: partial[slot] = the listing's projection over the tile.
template <int NB>
__global__ __launch_bounds__(256) void tile_gather(TorParams p, TileGeom g,
                                                   const TileChunk *__restrict__ chunks,
                                                   const int *__restrict__ list,
                                                   const float4 *__restrict__ rec, float *partial)
{
    extern __shared__ float lds[];
    const TileChunk ch = chunks[blockIdx.x];
    const TileBox b = tile_box(p, g, ch.tile);
    const int nv = g.S * g.Wy * g.Wx;
    for (int i = threadIdx.x; i < nv; i += blockDim.x) {
        int lx = i % g.Wx, r = i / g.Wx;
        int ly = r % g.Wy, ls = r / g.Wy;
        int ix = b.x0 + lx, iy = b.y0 + ly, s = b.sa + ls;
        float v = 0.0f;
        if (ix <= b.x1 && iy <= b.y1 && s <= b.sb)
            v = p.img[ix * p.sx + iy * p.sy + s * p.sz];
        lds[i] = v;
    }
    __syncthreads();
    const bool tof = p.use_tof;
    for (int j = threadIdx.x; j < ch.count; j += blockDim.x) {
        const int64_t e = list[ch.start + j];
        RawEvent r;
        Event ev;
        fetch_event_rec(rec, e, r);
        float acc = 0.0f;
        if (init_event<kProj>(p, r, ev))
            acc = tile_event<0, NB>(p, g, b, ev, tof, 0.0f, lds);
        partial[ch.start + j] = acc;
    }
}

// proj[e] = sum of the event's partials in its listing order.
__global__ __launch_bounds__(256) void tile_reduce(const float *__restrict__ partial,
                                                   const int *__restrict__ ev_off,
                                                   const int *__restrict__ ev_slot, int64_t n,
                                                   float *proj)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n)
        return;
    float v = 0.0f;
    for (int k = ev_off[e]; k < ev_off[e + 1]; k++)
        v += partial[ev_slot[k]];
    proj[e] = v;
}

} // namespace
} // namespace srfx

using namespace srfx;

struct srfx_tor_plan_s {
    TorParams p;  // geometry, model and events; image / value pointers set per call
    srfx_grid grid;
    TileGeom g;
    int cols7;
    int64_t nlist;
    int nchunks;
    int block;
    int *list;             // event of each slot, tiles in order
    float4 *rec;           // 48-byte record of each event
    TileChunk *chunks;
    float *proj;           // projection buffer of the OSEM step
    int *ev_off, *ev_slot; // tiled projection only: each event's slots
    float *partial;        // tiled projection only: per-slot partial sums
};

namespace {

template <class T> int dev_alloc(T **ptr, size_t count)
{
    *ptr = nullptr;
    if (count == 0)
        count = 1;
    if (hipMalloc((void **)ptr, sizeof(T) * count) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return SRFX_ENOMEM;
    }
    return SRFX_OK;
}

void plan_free(srfx_tor_plan_s *pl)
{
    if (!pl)
        return;
    (void)hipFree(pl->list);
    (void)hipFree(pl->rec);
    (void)hipFree(pl->chunks);
    (void)hipFree(pl->proj);
    (void)hipFree(pl->ev_off);
    (void)hipFree(pl->ev_slot);
    (void)hipFree(pl->partial);
    delete pl;
}

// LDS of the gather (FP32 tile) and scatter (FP64 tile) kernels.
unsigned lds_bytes(const TileGeom &g)
{
    return (unsigned)(sizeof(float) * g.S * g.Wx * g.Wy);
}

unsigned lds_bytes_scatter(const TileGeom &g)
{
    return (unsigned)(sizeof(double) * g.S * g.Wx * g.Wy);
}

// Largest tile count whose per-workgroup counters (3 per tile) fit in 64 KB.
constexpr int kLdsTiles = 16384 / 3;

template <bool FILL>
void launch_pass(const TorParams &p, const TileGeom &g, const PlanArgs &a, hipStream_t st)
{
    const dim3 grid(blocks_for(p.n, 256)), block(256);
    if (a.ntiles <= kLdsTiles)
        hipLaunchKernelGGL((plan_pass<FILL, true>), grid, block,
                           (FILL ? 3u : 1u) * sizeof(int) * (unsigned)a.ntiles, st, p, g, a);
    else
        hipLaunchKernelGGL((plan_pass<FILL, false>), grid, block, 0, st, p, g, a);
}

// Count, scan (host: per tile; device: per event when slots are kept), fill.
int plan_build(srfx_tor_plan_s *pl, hipStream_t st, int chunk, bool keep_slots)
{
    TorParams &p = pl->p;
    const TileGeom &g = pl->g;
    const int64_t n = p.n;
    const int ntiles = g.nslab * g.nwx * g.nwy;
    int *cnt_d = nullptr, *ev_cnt = nullptr, *tile_off_d = nullptr;
    void *tmp = nullptr;
    int rc = SRFX_OK;
    std::vector<int> cnt(ntiles), off(ntiles);
    std::vector<TileChunk> ch;
    int64_t total = 0;
    size_t tmp_bytes = 0;
    int last_cnt = 0, last_off = 0, mismatch = 0, total32 = 0;
    PlanArgs a{};
    // cnt_d: tile counts [ntiles], tile fill [ntiles], error count [1]
    if ((rc = dev_alloc(&cnt_d, 2 * (size_t)ntiles + 1)) || (rc = dev_alloc(&pl->rec, 3 * n)) ||
        (rc = dev_alloc(&tile_off_d, ntiles)) || (rc = dev_alloc(&pl->proj, n)))
        goto done;
    if (keep_slots && ((rc = dev_alloc(&ev_cnt, n)) || (rc = dev_alloc(&pl->ev_off, n + 1))))
        goto done;
    hipLaunchKernelGGL(fill_int, dim3(blocks_for(2 * (int64_t)ntiles + 1, 256)), dim3(256), 0,
                       st, cnt_d, 2 * (int64_t)ntiles + 1, 0);
    a.cols7 = pl->cols7;
    a.ntiles = ntiles;
    a.tile_cnt = cnt_d;
    a.ev_cnt = ev_cnt;
    a.rec = pl->rec;
    if (n > 0)
        launch_pass<false>(p, g, a, st);
    if ((rc = launch_status()))
        goto done;
    if (keep_slots && n > 0) {
        // ev_off = exclusive scan of the per-event listing counts
        if (rocprim::exclusive_scan(nullptr, tmp_bytes, ev_cnt, pl->ev_off, 0, (size_t)n,
                                    rocprim::plus<int>(), st) != hipSuccess) {
            rc = SRFX_ELAUNCH;
            goto done;
        }
        if ((rc = dev_alloc((char **)&tmp, tmp_bytes)))
            goto done;
        if (rocprim::exclusive_scan(tmp, tmp_bytes, ev_cnt, pl->ev_off, 0, (size_t)n,
                                    rocprim::plus<int>(), st) != hipSuccess ||
            hipMemcpyAsync(&last_cnt, ev_cnt + (n - 1), sizeof(int), hipMemcpyDeviceToHost,
                           st) != hipSuccess ||
            hipMemcpyAsync(&last_off, pl->ev_off + (n - 1), sizeof(int), hipMemcpyDeviceToHost,
                           st) != hipSuccess) {
            rc = SRFX_ELAUNCH;
            goto done;
        }
    }
    if (hipMemcpyAsync(cnt.data(), cnt_d, sizeof(int) * ntiles, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    for (int t = 0; t < ntiles; t++) {
        off[t] = (int)total;
        total += cnt[t];
        if (total > (int64_t)INT32_MAX - 1)
            break;
    }
    if (total > (int64_t)INT32_MAX - 1 ||
        (keep_slots && n > 0 && (int64_t)last_off + last_cnt != total)) {
        rc = SRFX_EUNSUPPORTED;
        goto done;
    }
    total32 = (int)total;
    for (int t = 0; t < ntiles; t++)
        for (int s = 0; s < cnt[t]; s += chunk)
            ch.push_back(TileChunk{t, off[t] + s, cnt[t] - s < chunk ? cnt[t] - s : chunk, 0});
    pl->nlist = total;
    pl->nchunks = (int)ch.size();
    if ((rc = dev_alloc(&pl->list, total)) || (rc = dev_alloc(&pl->chunks, ch.size())))
        goto done;
    if (keep_slots && ((rc = dev_alloc(&pl->ev_slot, total)) ||
                       (rc = dev_alloc(&pl->partial, total))))
        goto done;
    if (hipMemcpyAsync(tile_off_d, off.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        (!ch.empty() && hipMemcpyAsync(pl->chunks, ch.data(), sizeof(TileChunk) * ch.size(),
                                       hipMemcpyHostToDevice, st) != hipSuccess) ||
        (keep_slots && hipMemcpyAsync(pl->ev_off + n, &total32, sizeof(int),
                                      hipMemcpyHostToDevice, st) != hipSuccess)) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    a.ev_cnt = nullptr;
    a.ev_off = keep_slots ? pl->ev_off : nullptr;
    a.tile_off = tile_off_d;
    a.tile_fill = cnt_d + ntiles;
    a.list = pl->list;
    a.ev_slot = keep_slots ? pl->ev_slot : nullptr;
    a.err = cnt_d + 2 * ntiles;
    if (n > 0)
        launch_pass<true>(p, g, a, st);
    if ((rc = launch_status()))
        goto done;
    if (hipMemcpyAsync(&mismatch, cnt_d + 2 * ntiles, sizeof(int), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        rc = SRFX_ELAUNCH;
        goto done;
    }
    if (mismatch)
        rc = SRFX_ELAUNCH; // count and fill passes disagreed: no valid plan
done:
    // host buffers above are read by the copies: finish them before returning
    (void)hipStreamSynchronize(st);
    (void)hipFree(cnt_d);
    (void)hipFree(ev_cnt);
    (void)hipFree(tile_off_d);
    (void)hipFree(tmp);
    return rc;
}

template <int MODE>
void launch_scatter(const TorParams &p, const srfx_tor_plan_s *pl, dim3 grid, dim3 block,
                    unsigned sh, hipStream_t st)
{
    switch (p.box) {
    case 3:
        hipLaunchKernelGGL((tile_scatter<MODE, 3>), grid, block, sh, st, p, pl->g, pl->chunks,
                           pl->list, pl->rec);
        break;
    case 4:
        hipLaunchKernelGGL((tile_scatter<MODE, 4>), grid, block, sh, st, p, pl->g, pl->chunks,
                           pl->list, pl->rec);
        break;
    default:
        hipLaunchKernelGGL((tile_scatter<MODE, 0>), grid, block, sh, st, p, pl->g, pl->chunks,
                           pl->list, pl->rec);
        break;
    }
}

int plan_scatter(srfx_tor_plan_s *pl, int mode, const float *vals, float *image,
                 const srfx_grid *out_grid, uint32_t flags, hipStream_t st)
{
    TorParams p = pl->p;
    p.val = vals;
    p.out = image;
    if (out_grid) {
        Strides so = grid_strides(*out_grid);
        p.ox = (int)so.x;
        p.oy = (int)so.y;
        p.oz = (int)so.z;
    }
    if (pl->nchunks == 0)
        return SRFX_OK;
    const dim3 grid(pl->nchunks), block(pl->block);
    const unsigned sh = lds_bytes_scatter(pl->g);
    switch (mode) {
    case kBack:
        launch_scatter<kBack>(p, pl, grid, block, sh, st);
        break;
    case kAdj:
        launch_scatter<kAdj>(p, pl, grid, block, sh, st);
        break;
    default:
        launch_scatter<kMap>(p, pl, grid, block, sh, st);
        break;
    }
    (void)flags;
    return launch_status();
}

int plan_project(srfx_tor_plan_s *pl, const float *image, float *proj, hipStream_t st)
{
    TorParams p = pl->p;
    p.img = image;
    if (!pl->ev_slot)
        return SRFX_EUNSUPPORTED; // plan built without the tiled projection tables
    if (p.n == 0)
        return SRFX_OK;
    if (pl->nchunks > 0) {
        const dim3 grid(pl->nchunks), block(pl->block);
        const unsigned sh = lds_bytes(pl->g);
        switch (p.box) {
        case 3:
            hipLaunchKernelGGL(tile_gather<3>, grid, block, sh, st, p, pl->g, pl->chunks,
                               pl->list, pl->rec, pl->partial);
            break;
        case 4:
            hipLaunchKernelGGL(tile_gather<4>, grid, block, sh, st, p, pl->g, pl->chunks,
                               pl->list, pl->rec, pl->partial);
            break;
        default:
            hipLaunchKernelGGL(tile_gather<0>, grid, block, sh, st, p, pl->g, pl->chunks,
                               pl->list, pl->rec, pl->partial);
            break;
        }
    }
    hipLaunchKernelGGL(tile_reduce, dim3(blocks_for(p.n, 256)), dim3(256), 0, st, pl->partial,
                       pl->ev_off, pl->ev_slot, p.n, proj);
    return launch_status();
}

} // namespace

extern "C" int srfx_tor_plan_create(const float *lors, int64_t n, int64_t ld, int32_t ncols,
                                    srfx_grid grid, srfx_tor_model m, srfx_tile tile,
                                    srfx_tor_plan *out, void *stream)
{
    if (!out)
        return SRFX_EINVAL;
    *out = nullptr;
    if (ncols != 7 && ncols != 10)
        return SRFX_EINVAL;
    if (n > (int64_t)INT32_MAX - 1)
        return SRFX_EUNSUPPORTED;
    if (tile.slices < 1 || tile.wx < 1 || tile.wy < 1 || tile.chunk < 1)
        return SRFX_EINVAL;
    if ((int64_t)tile.slices * tile.wx * tile.wy > 8192)
        return SRFX_EUNSUPPORTED; // FP64 LDS tile above 64 KB
    if (tile.block != 0 && tile.block != 64 && tile.block != 128 && tile.block != 256)
        return SRFX_EINVAL;
    srfx_tor_plan_s *pl = new srfx_tor_plan_s{};
    srfx_tor_model mm = ncols == 7 ? model_no_tof(m.kernel_width, m.geo_sigma2_flag) : m;
    int rc = setup(pl->p, lors, n, ld, grid, mm.kernel_width, mm.geo_sigma2_flag, mm.tof_bin,
                   mm.tof_sigma2);
    if (rc) {
        delete pl;
        return rc;
    }
    pl->grid = grid;
    pl->cols7 = ncols == 7;
    pl->block = tile.block ? tile.block : 256;
    TileGeom &g = pl->g;
    g.S = tile.slices;
    g.Wx = tile.wx;
    g.Wy = tile.wy;
    g.nwx = (pl->p.gx + g.Wx - 1) / g.Wx;
    g.nwy = (pl->p.gy + g.Wy - 1) / g.Wy;
    g.nslab = (pl->p.gz + g.S - 1) / g.S;
    if ((int64_t)g.nwx * g.nwy * g.nslab > (int64_t)INT32_MAX / 2) {
        delete pl;
        return SRFX_EUNSUPPORTED;
    }
    rc = plan_build(pl, (hipStream_t)stream, tile.chunk, tile.projection != 0);
    if (rc) {
        plan_free(pl);
        return rc;
    }
    *out = pl;
    return SRFX_OK;
}

extern "C" int srfx_tor_plan_info(srfx_tor_plan plan, int64_t *nlistings, int32_t *nchunks)
{
    if (!plan)
        return SRFX_EINVAL;
    if (nlistings)
        *nlistings = plan->nlist;
    if (nchunks)
        *nchunks = plan->nchunks;
    return SRFX_OK;
}

extern "C" int srfx_tor_plan_destroy(srfx_tor_plan plan)
{
    if (plan)
        (void)hipDeviceSynchronize();
    plan_free(plan);
    return SRFX_OK;
}

extern "C" int srfx_tor_plan_project(srfx_tor_plan plan, const float *image, float *proj,
                                     uint32_t flags, void *stream)
{
    if (!plan || plan->cols7 || (plan->p.n > 0 && (!image || !proj)))
        return SRFX_EINVAL;
    (void)flags;
    return plan_project(plan, image, proj, (hipStream_t)stream);
}

static int plan_scatter_entry(srfx_tor_plan plan, int mode, const float *vals, float *image,
                              uint32_t flags, void *stream)
{
    if (!plan || !image || (plan->p.n > 0 && !vals))
        return SRFX_EINVAL;
    if ((mode == kMap) != (plan->cols7 != 0))
        return SRFX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    int rc = zero_grid(image, plan->grid, flags, st);
    if (rc)
        return rc;
    return plan_scatter(plan, mode, vals, image, nullptr, flags, st);
}

extern "C" int srfx_tor_plan_backproject(srfx_tor_plan plan, const float *lors_value,
                                         float *image, uint32_t flags, void *stream)
{
    return plan_scatter_entry(plan, kBack, lors_value, image, flags, stream);
}

extern "C" int srfx_tor_plan_backproject_adjoint(srfx_tor_plan plan, const float *g,
                                                 float *image, uint32_t flags, void *stream)
{
    return plan_scatter_entry(plan, kAdj, g, image, flags, stream);
}

extern "C" int srfx_tor_plan_maplors(srfx_tor_plan plan, const float *lors_value, float *image,
                                     uint32_t flags, void *stream)
{
    return plan_scatter_entry(plan, kMap, lors_value, image, flags, stream);
}

extern "C" int srfx_tor_plan_osem_subset(srfx_tor_plan plan, const float *image,
                                         srfx_grid bp_grid, float *bp, uint32_t flags,
                                         void *stream)
{
    if (!plan || plan->cols7 || !bp || (plan->p.n > 0 && !image) || !grid_ok(bp_grid))
        return SRFX_EINVAL;
    const TorParams &p = plan->p;
    if (bp_grid.shape[0] != p.gx || bp_grid.shape[1] != p.gy || bp_grid.shape[2] != p.gz)
        return SRFX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    int rc = zero_grid(bp, bp_grid, flags, st);
    if (rc)
        return rc;
    if (p.n == 0)
        return SRFX_OK;
    // Projection with the chosen schedule (untiled by default), then the
    // tiled scatter of 1 / p over events with p > 1e-7.
    if (flags & SRFX_TILED_PROJECTION) {
        rc = plan_project(plan, image, plan->proj, st);
    } else {
        TorParams q = p;
        q.img = image;
        q.out = plan->proj;
        rc = launch<kProj>(q, flags, st);
    }
    if (rc)
        return rc;
    return plan_scatter(plan, kBack, plan->proj, bp, &bp_grid, flags, st);
}
