// Device side of the tiled TOR ops (tile.hip: plans and their entry points;
// batch.hip: batched reconstructions over plans), and the plan object.
//
// A plan bins every event of a fixed set into the image tiles it can touch;
// the kernels then gather from / scatter into tiles staged in LDS.
//
// A tile is a slab of S consecutive slices times a window of Wx x Wy voxels
// of the op frame.  Along a TOR group's dominant axis an event moves by at
// most one voxel pitch per slice, so within one slab its voxels stay inside a
// small footprint and it touches few tiles.  The plan lists, for every tile,
// the events whose footprint meets it ("listings"), cut into chunks of at most
// `chunk` listings; one workgroup takes one chunk:
// Each event's walk record (tor_core.h) is computed once when the plan is
// built; a listing unpacks it instead of redoing the event's set-up.
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
#pragma once

#include <atomic>
#include <vector>

#include "tor_core.h"

namespace srfx {

struct TileGeom {
    int S, Wx, Wy;       // slices per slab, window width along op x and y (voxels)
    int nwx, nwy, nslab; // windows per axis and number of slabs
    int sorted;          // chunks' listings ordered by slice count (sort_chunks)
    int spans;           // list entries carry their listing's slices (kSpanEventBits)
};

// A list entry of a plan with spans: the event index in bits 0..23 and the
// listing's first and last slice in the tile's slab, relative to the slab's
// first slice, in bits 24..27 and 28..31 (slabs of at most 16 slices, plans
// of fewer than 2^24 events).  An empty span is stored as (1, 0).
constexpr int kSpanEventBits = 24;
constexpr int kSpanMaxSlab = 16;


struct TileChunk {
    int tile, start, count, pad;
};

namespace {

__device__ __forceinline__ int trunc_clamped(double u)
{
    return (int)fmin(fmax(u, -1.0e9), 1.0e9);
}

// Slack (voxels) on the candidate extents below: the walks form their
// offsets in FP32, ~1e-6 voxel from the FP64 track.
constexpr double kExtentSlack = 1.0e-3;

// Half extents, in voxels, of the padded tube ellipse of a slice around the
// crossing point (u, w): a voxel the walks keep (d2 < 9 sigma_c^2) or decide
// in FP64 (within 1e-4 mm^2 of the cut-off) has its centre inside that
// ellipse (cut-off tm, make_track), so its column i and row j satisfy
// |i + 0.5 - u| <= hx / inter_x and |j + 0.5 - w| <= dymax / inter_y.
// Voxels of the candidate box outside the ellipse add nothing.
__device__ __forceinline__ double col_extent(const TorParams &p, const Track &t)
{
    return (double)t.hx * p.rinter_x + kExtentSlack;
}
__device__ __forceinline__ double row_extent(const TorParams &p, const Track &t)
{
    return (double)t.dymax * p.rinter_y + kExtentSlack;
}

// Windows wx0..wx1, wy0..wy1 holding every voxel the event can visit in the
// slices s0..s1: the image's columns and rows that the padded ellipse covers
// over the slab (u and w are affine in s).  Every walk adds only to voxels
// whose centre lies inside that ellipse, whatever the patch origin.  (The
// walks also stay inside the patch; intersecting with its span as well
// listed no fewer tiles at 2 mm pitch for kernel widths 4 and 8 mm.)
__device__ __forceinline__ bool footprint(const TorParams &p, const TileGeom &g, const Track &t,
                                          int s0, int s1, int &wx0, int &wx1, int &wy0, int &wy1)
{
    const double u0 = t.ax + (double)s0 * t.bx, u1 = t.ax + (double)s1 * t.bx;
    const double w0 = t.ay + (double)s0 * t.by, w1 = t.ay + (double)s1 * t.by;
    const double ex = col_extent(p, t), ey = row_extent(p, t);
    const int x0 = max(0, trunc_clamped(ceil(fmin(u0, u1) - 0.5 - ex)));
    const int x1 = min(p.gx - 1, trunc_clamped(floor(fmax(u0, u1) - 0.5 + ex)));
    const int y0 = max(0, trunc_clamped(ceil(fmin(w0, w1) - 0.5 - ey)));
    const int y1 = min(p.gy - 1, trunc_clamped(floor(fmax(w0, w1) - 0.5 + ey)));
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
                f((sl * g.nwy + wy) * g.nwx + wx, sl, wx, wy);
    }
}

} // namespace

struct TileBox {
    int sa, sb, x0, x1, y0, y1;
};

namespace {

// Tile bounds of slab sl, window (wx, wy): slices sa..sb, op-frame voxels
// x0..x1, y0..y1.
__device__ __forceinline__ TileBox tile_box_at(const TorParams &p, const TileGeom &g, int sl,
                                               int wx, int wy)
{
    TileBox b;
    b.sa = sl * g.S;
    b.sb = min(b.sa + g.S, p.gz) - 1;
    b.x0 = wx * g.Wx;
    b.x1 = min(b.x0 + g.Wx, p.gx) - 1;
    b.y0 = wy * g.Wy;
    b.y1 = min(b.y0 + g.Wy, p.gy) - 1;
    return b;
}

// Tile bounds of a chunk's tile index.
__device__ __forceinline__ TileBox tile_box(const TorParams &p, const TileGeom &g, int tile)
{
    const int per = g.nwx * g.nwy;
    const int sl = tile / per, r = tile - sl * per;
    const int wy = r / g.nwx, wx = r - wy * g.nwx;
    return tile_box_at(p, g, sl, wx, wy);
}

// Whether a scatter into p.out should use the y-fastest LDS tile.
__host__ __device__ inline bool flush_y_fastest(const TorParams &p)
{
    return (p.oy < 0 ? -p.oy : p.oy) < (p.ox < 0 ? -p.ox : p.ox);
}

// Index of voxel (ix, iy) of slice s in the LDS tile of box b: x fastest, or
// (YF) y fastest, so that a flush along the tile's fastest axis writes
// consecutive addresses of an output whose unit stride is along op y.

// Narrow [s0, s1] to the slices s with lo <= u(s) <= hi, u(s) = a + s b,
// rb = 1 / b (FP64; the bounds on s are widened by 1e-6 slice, so the result
// is a superset).
__device__ __forceinline__ void axis_slices(double a, double b, double rb, double lo, double hi,
                                            int &s0, int &s1)
{
    if (fabs(b) < 1.0e-12) {
        // u is constant to 1e-12 voxel per slice: test it at both ends
        const double u0 = a + (double)s0 * b, u1 = a + (double)s1 * b;
        if (fmax(u0, u1) < lo - 1.0e-6 || fmin(u0, u1) > hi + 1.0e-6)
            s1 = s0 - 1;
        return;
    }
    const double t1 = (lo - a) * rb, t2 = (hi - a) * rb;
    const double tl = fmax(fmin(t1, t2), -1.0e8), th = fmin(fmax(t1, t2), 1.0e8);
    s0 = max(s0, (int)ceil(tl - 1.0e-6));
    s1 = min(s1, (int)floor(th + 1.0e-6));
}

// 1 / b for axis_slices (any value where |b| < 1e-12: not used there).
__device__ __forceinline__ double slope_rcp(double b)
{
    return fabs(b) < 1.0e-12 ? 0.0 : 1.0 / b;
}

// Narrow [s0, s1] to a superset of the slices at which the walk can add to
// a voxel of columns x0..x1 and rows y0..y1: u(s) within the column extent
// of the window, [x0 + 0.5 - hx', x1 + 0.5 + hx'], and w(s) within the row
// extent (every walk adds only to voxels inside the padded ellipse).  The
// walks also stay inside the patch, but its condition on u, [x0 - P + half
// - 2, x1 + half + 2] with 2 voxels of slack for the exact origin, follows
// from the column extent's whenever hx' < min(half, P - half) + 1.5 voxels,
// as for the candidate-box events; it is not tested.  rbx, rby: slope_rcp
// of the track's slopes.
__device__ __forceinline__ void window_slices(const TorParams &p, const Track &t, double rbx,
                                              double rby, int x0, int x1, int y0, int y1, int &s0,
                                              int &s1)
{
    const double ex = col_extent(p, t), ey = row_extent(p, t);
    axis_slices(t.ax, t.bx, rbx, x0 + 0.5 - ex, x1 + 0.5 + ex, s0, s1);
    axis_slices(t.ay, t.by, rby, y0 + 0.5 - ey, y1 + 0.5 + ey, s0, s1);
}

// The slices s0..s1 a listing walks: the event's crossed slices in the
// tile's slab, narrowed to those at which the walk can visit the window.
__device__ __forceinline__ void listing_slices(const TorParams &p, const TileBox &b,
                                               const Event &ev, const Track &t, double rbx,
                                               double rby, int &s0, int &s1)
{
    s0 = max(ev.s_lo, b.sa);
    s1 = min(ev.s_hi, b.sb);
    window_slices(p, t, rbx, rby, b.x0, b.x1, b.y0, b.y1, s0, s1);
}

__device__ __forceinline__ void listing_slices(const TorParams &p, const TileBox &b,
                                               const Event &ev, const Track &t, int &s0, int &s1)
{
    listing_slices(p, b, ev, t, slope_rcp(t.bx), slope_rcp(t.by), s0, s1);
}

struct PlanArgs {
    int cols7, ntiles;
    int *tile_cnt;        // count pass: listings per tile; fill pass: read
    int *ev_cnt;          // count pass: listings per event (nullable)
    float4 *rec;          // count pass writes, fill pass reads: walk records
    const int *ev_off;    // fill pass: first listing of each event (nullable)
    const int *tile_off;  // fill pass: first slot of each tile
    int *tile_fill;       // fill pass: slots taken per tile
    int *list;            // fill pass: event of each slot
    int *ev_slot;         // fill pass: slot of each of an event's listings (nullable)
    int *err;             // fill pass: count of listings that did not fit
    int *other;           // count pass: set to 1 if an event needs another walk than
                          // the NB x (NB - 1) box (nullable)
};

// Whether the walks of an event take the NB x (NB - 1) candidate box (box
// shape 1) in every slice: then a plan's scatter can run the kernel that
// holds only that walk.
__device__ __forceinline__ bool shape_one(const TorParams &p, const Track &t)
{
    return !p.full_patch && p.box > 0 && t.nx <= p.box && t.ny <= p.box - 1;
}

// Set *flag once an event needs another walk: a relaxed load first, so that
// once the flag is visible the lanes only read it (one cached load per wave)
// instead of all storing to one address.
__device__ __forceinline__ void flag_other(int *flag)
{
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool plan_init(const TorParams &p, int cols7, const RawEvent &r,
                                          const ExactGeom &eg, Event &ev)
{
    return cols7 ? init_event<kMap>(p, r, eg, ev) : init_event<kProj>(p, r, eg, ev);
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
    Event ev{};
    Track t{};
    bool ok = false;
    if (valid) {
        if (FILL) {
            // the walk record the count pass wrote
            unpack_walk(p, a.rec, e, ev, t);
            ok = ev.s_lo <= ev.s_hi;
        } else {
            RawEvent r;
            if (a.cols7)
                fetch_event<kMap>(p, e, r);
            else
                fetch_event<kProj>(p, e, r);
            const ExactGeom eg = exact_geom(p, r);
            ok = plan_init(p, a.cols7, r, eg, ev);
            if (ok)
                make_track(p, ev, eg, r, t);
            pack_walk(p, r, ev, t, ok, a.rec, e);
        }
    }
    if (!FILL) {
        int k = 0;
        if (ok && a.other && p.box > 0 && !shape_one(p, t))
            flag_other(a.other);
        if (ok)
            for_each_tile(p, g, ev, t, [&](int tile, int, int, int) {
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
            for_each_tile(p, g, ev, t, [&](int tile, int, int, int) { atomicAdd(&sh[tile], 1); });
        __syncthreads();
        for (int i = threadIdx.x; i < nt; i += blockDim.x) {
            int c = sh[i];
            sh[nt + i] = c ? atomicAdd(&a.tile_fill[i], c) : 0;
        }
        __syncthreads();
    }
    const int k0 = (valid && a.ev_slot) ? a.ev_off[e] : 0;
    const int k1 = (valid && a.ev_slot) ? a.ev_off[e + 1] : 0;
    const double rbx = slope_rcp(t.bx), rby = slope_rcp(t.by);
    int k = 0;
    if (ok)
        for_each_tile(p, g, ev, t, [&](int tile, int sl, int wx, int wy) {
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
                int ent = (int)e;
                if (g.spans) {
                    // the listing's slices ride in the entry (the host checked
                    // the limits)
                    const TileBox b = tile_box_at(p, g, sl, wx, wy);
                    int s0, s1;
                    listing_slices(p, b, ev, t, rbx, rby, s0, s1);
                    const int lo = s0 <= s1 ? s0 - b.sa : 1, hi = s0 <= s1 ? s1 - b.sa : 0;
                    ent |= (lo << kSpanEventBits) | (hi << (kSpanEventBits + 4));
                }
                a.list[slot] = ent;
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

// A set of plans over consecutive event ranges of one LOR array (one plan per
// OSEM subset of an axis group): plan k holds events [first[k], first[k+1]),
// indexed from 0 in its own records and list.  Tables are per (plan, tile),
// plan-major.
struct PlanSetArgs {
    int cols7, ntiles, nplans;
    int lds;                // the per-workgroup counters fit in LDS
    int queue;              // fill pass: LDS also holds listing queues of this many
                            // entries per wave (0: no queues)
    const int64_t *first;   // [nplans + 1]
    float4 *const *rec;     // [nplans] walk records
    int *cnt;               // count pass: listings per (plan, tile); fill pass: read
    const int *off;         // fill pass: first slot of each (plan, tile) in the plan's list
    int *fill;              // fill pass: slots taken per (plan, tile)
    int *const *list;       // fill pass: [nplans] list entries
    int *err;               // fill pass: listings that did not fit
    int *other;             // count pass: per plan, set to 1 if an event needs
                            // another walk than the NB x (NB - 1) box
};

// The plan holding event e of the set's array.
__device__ __forceinline__ int plan_of(const PlanSetArgs &a, int64_t e)
{
    int lo = 0, hi = a.nplans; // first[lo] <= e < first[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.first[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// The fill pass's listing queues: per wave, up to kFillQueue entries (tile |
// the event's lane << 24) in LDS, and per event what its listings' slices
// need.  A queue shorter than kFillQueueMin would overflow too often (a
// wave lists ~700 (event, tile) pairs at 11 per event).
constexpr int kFillQueue = 1280, kFillQueueMin = 768;
struct FillEvent {
    double ax, bx, ay, by, rbx, rby;
    float hx, dymax;
    int s_lo, s_hi;
};
// LDS bytes of the fill pass with queues of qcap entries, ahead of its
// three counter arrays
__host__ __device__ constexpr unsigned fill_queue_lds(int qcap)
{
    return 256u * sizeof(FillEvent) + 4u * (unsigned)qcap * sizeof(int) + 4u * sizeof(int);
}

// listing_slices from a queued event's fields
__device__ __forceinline__ void listing_slices(const TorParams &p, const TileBox &b,
                                               const FillEvent &f, int &s0, int &s1)
{
    Track t;
    t.ax = f.ax;
    t.bx = f.bx;
    t.ay = f.ay;
    t.by = f.by;
    t.hx = f.hx;
    t.dymax = f.dymax;
    s0 = max(f.s_lo, b.sa);
    s1 = min(f.s_hi, b.sb);
    window_slices(p, t, f.rbx, f.rby, b.x0, b.x1, b.y0, b.y1, s0, s1);
}

// plan_pass for a plan set, over all events of the set in one launch.  A
// workgroup whose events all belong to one plan counts in LDS, as plan_pass
// does for one plan; a workgroup across a plan boundary (one per boundary)
// uses the global counters directly.  List entries carry the listing's
// slices (the set's plans always have spans).
//
// The fill pass with queues (a.queue, workgroups of 256 counting in LDS):
// each lane walks its event's tiles once, counting them in LDS and queueing
// (tile, event) in its wave's queue; after the workgroup has reserved its
// slots, the 64 lanes of a wave take the wave's queued listings in turn
// (their slices and slots), so that a wave's work follows the listings of
// its 64 events instead of its busiest event.  Listings past a full queue
// are listed by their own lane afterwards, walking its tiles again.
template <bool FILL>
__global__ __launch_bounds__(256) void plan_set_pass(TorParams p, TileGeom g, PlanSetArgs a)
{
    // LDS: with queues, [FillEvent x 256][4 queues of a.queue entries][4
    // queue lengths] first; then [counts (FILL: + bases, ranks)][ntiles]
    extern __shared__ int4 sh4[];
    const bool queued = FILL && a.queue > 0;
    const int qcap = a.queue;
    char *const shb = reinterpret_cast<char *>(sh4);
    int *const sh = reinterpret_cast<int *>(shb + (queued ? fill_queue_lds(qcap) : 0u));
    FillEvent *const fev = reinterpret_cast<FillEvent *>(shb);
    int *const queue = reinterpret_cast<int *>(shb + 256u * sizeof(FillEvent));
    int *const qlen = queue + 4 * qcap;
    const int nt = a.ntiles;
    const int per = FILL ? 3 * nt : nt;
    const int64_t total = a.first[a.nplans];
    const int64_t e0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t e = e0 + threadIdx.x;
    const bool valid = e < total;
    const int kf = plan_of(a, e0), kl = plan_of(a, min(e0 + (int64_t)blockDim.x, total) - 1);
    const bool lds = a.lds && kl == kf;
    const bool useq = queued && lds;
    const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    if (lds) {
        for (int i = threadIdx.x; i < per; i += blockDim.x)
            sh[i] = 0;
        if (useq && threadIdx.x < 4)
            qlen[threadIdx.x] = 0;
        __syncthreads();
    }
    const int k = valid ? plan_of(a, e) : kf;
    const int64_t le = e - a.first[k];
    TorParams q = p; // plan k's events
    q.lors = p.lors + a.first[k];
    q.n = a.first[k + 1] - a.first[k];
    int *lcnt = sh;
    int *gcnt = a.cnt + (int64_t)k * nt;
    Event ev{};
    Track t{};
    bool ok = false;
    if (valid) {
        if (FILL) {
            unpack_walk(q, a.rec[k], le, ev, t);
            ok = ev.s_lo <= ev.s_hi;
        } else {
            RawEvent r;
            if (a.cols7)
                fetch_event<kMap>(q, le, r);
            else
                fetch_event<kProj>(q, le, r);
            const ExactGeom eg = exact_geom(q, r);
            ok = plan_init(q, a.cols7, r, eg, ev);
            if (ok)
                make_track(q, ev, eg, r, t);
            pack_walk(q, r, ev, t, ok, a.rec[k], le);
            if (ok && q.box > 0 && !shape_one(q, t))
                flag_other(&a.other[k]);
        }
    }
    int drop = -1; // with queues: this lane's first listing past its wave's full queue
    if (useq) {
        if (ok) {
            fev[threadIdx.x] = FillEvent{t.ax, t.bx, t.ay, t.by, slope_rcp(t.bx),
                                         slope_rcp(t.by), t.hx, t.dymax, ev.s_lo, ev.s_hi};
            int kk = 0;
            for_each_tile(q, g, ev, t, [&](int tile, int, int, int) {
                atomicAdd(&lcnt[tile], 1);
                // the active lanes take consecutive queue slots: one LDS
                // atomic per wave, by the lowest of them
                const uint64_t m = __ballot(1);
                const int r = (int)__builtin_amdgcn_mbcnt_hi(
                    (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                int base = 0;
                if (r == 0)
                    base = atomicAdd(&qlen[wv], __popcll(m));
                base = __shfl(base, __builtin_ctzll(m), 64);
                if (base + r < qcap)
                    queue[wv * qcap + base + r] = tile | (lane << 24);
                else if (drop < 0)
                    drop = kk;
                kk++;
            });
        }
    } else if (!FILL || lds) {
        if (ok)
            for_each_tile(q, g, ev, t, [&](int tile, int, int, int) {
                atomicAdd(lds ? &lcnt[tile] : &gcnt[tile], 1);
            });
    }
    if (lds) {
        __syncthreads();
        for (int tile = threadIdx.x; tile < nt; tile += blockDim.x) {
            const int c = sh[tile];
            if (c) {
                if (FILL)
                    sh[nt + tile] = atomicAdd(&a.fill[(int64_t)kf * nt + tile], c);
                else
                    atomicAdd(&a.cnt[(int64_t)kf * nt + tile], c);
            }
        }
        __syncthreads();
    }
    if (!FILL)
        return;
    const int *goff = a.off + (int64_t)k * nt;
    int *gfill = a.fill + (int64_t)k * nt;
    int *list = a.list[k];
    // one listing of a workgroup counting in LDS: its slot, its entry
    auto put = [&](int tile, const TileBox &b, int s0, int s1, int64_t lev) {
        const int loc = atomicAdd(&lcnt[2 * nt + tile], 1);
        const int rank = lcnt[nt + tile] + loc;
        if (loc < lcnt[tile] && rank < gcnt[tile]) {
            const int lo = s0 <= s1 ? s0 - b.sa : 1, hi = s0 <= s1 ? s1 - b.sa : 0;
            list[goff[tile] + rank] =
                (int)lev | (lo << kSpanEventBits) | (hi << (kSpanEventBits + 4));
        } else {
            atomicAdd(a.err, 1);
        }
    };
    if (useq) {
        // every valid lane of the workgroup is in plan kf
        const int nq = min(qlen[wv], qcap);
        const int *wq = queue + wv * qcap;
        const int64_t le0 = e0 + 64 * wv - a.first[kf];
        for (int i = lane; i < nq; i += 64) {
            const int ent = wq[i];
            const int tile = ent & 0xffffff, el = (int)((unsigned)ent >> 24);
            const FillEvent f = fev[64 * wv + el];
            const TileBox b = tile_box(q, g, tile);
            int s0, s1;
            listing_slices(q, b, f, s0, s1);
            put(tile, b, s0, s1, le0 + el);
        }
        if (drop >= 0) {
            const double rbx = slope_rcp(t.bx), rby = slope_rcp(t.by);
            int kk = 0;
            for_each_tile(q, g, ev, t, [&](int tile, int sl, int wx, int wy) {
                if (kk++ >= drop) {
                    const TileBox b = tile_box_at(q, g, sl, wx, wy);
                    int s0, s1;
                    listing_slices(q, b, ev, t, rbx, rby, s0, s1);
                    put(tile, b, s0, s1, le);
                }
            });
        }
        return;
    }
    const double rbx = slope_rcp(t.bx), rby = slope_rcp(t.by);
    if (ok)
        for_each_tile(q, g, ev, t, [&](int tile, int sl, int wx, int wy) {
            const TileBox b = tile_box_at(q, g, sl, wx, wy);
            if (lds) {
                int s0, s1;
                listing_slices(q, b, ev, t, rbx, rby, s0, s1);
                put(tile, b, s0, s1, le);
                return;
            }
            const int rank = atomicAdd(&gfill[tile], 1);
            if (rank < gcnt[tile]) {
                int s0, s1;
                listing_slices(q, b, ev, t, rbx, rby, s0, s1);
                const int lo = s0 <= s1 ? s0 - b.sa : 1, hi = s0 <= s1 ? s1 - b.sa : 0;
                list[goff[tile] + rank] =
                    (int)le | (lo << kSpanEventBits) | (hi << (kSpanEventBits + 4));
            } else {
                atomicAdd(a.err, 1);
            }
        });
}

__global__ __launch_bounds__(256) void fill_int(int *v, int64_t n, int value)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        v[i] = value;
}

template <bool YF>
__device__ __forceinline__ int lds_index(const TileGeom &g, const TileBox &b, int s, int ix, int iy)
{
    return YF ? ((s - b.sa) * g.Wx + (ix - b.x0)) * g.Wy + (iy - b.y0)
              : ((s - b.sa) * g.Wy + (iy - b.y0)) * g.Wx + (ix - b.x0);
}

// The chord walk of one slice (patch_walk) clipped to the tile's window.
//   KIND 0: returns sum w * lds[voxel]; KIND 1: lds[voxel] += w (amp carries
//   the scale).
template <int KIND, bool YF, typename T>
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
    const int org = lds_index<YF>(g, b, s, c.ix0, c.iy0);
    const int si = YF ? g.Wy : 1, sj = YF ? 1 : g.Wx;
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
            const float dt = voxel_dt(ev, dx, dy);
            bool in = dt < -kThrMargin;
            if (!in && dt <= kThrMargin)
                in = inside_exact(p, ev, s, c.ix0 + i, c.iy0 + j);
            if (in) {
                const float w = voxel_weight(ev, amp, dt);
                int li = org + j * sj + i * si;
                if constexpr (KIND == 0)
                    acc = fmaf(w, lds[li], acc);
                else
                    atomicAdd(&lds[li], (T)w);
            }
        }
    }
    return acc;
}

// One listing: its slices s0..s1 inside the tile (listing_slices), each
// through the NB x NB candidate box clipped to the window when the event's
// box fits, else the clipped chord walk.
// The scatter of one listing (slices s0..s1) through its BX x BY candidate box
// into a y-fastest FP64 LDS tile: the box walk of slice_box<1, BX, BY, 2>
// with less work per slice.
//   * The box's clip to the patch [0, P) and the window as float bounds
//     (lo, hi) on its first column / row, so the offsets need no integer
//     round trip; clipped candidates get an infinite term (neither kept nor
//     near).
//   * The scale, the TOF constant and the slice's TOF factor folded into each
//     kept candidate's exponent, exp2(cw dt + thr cw + log2 |scale tofk| +
//     TOF quadratic), times the sign of the scale unless POS (a positive
//     scale: the ratio backprojection).
//   * LDS byte offsets carried from slice to slice.
// The same candidates and decisions as slice_box (its box contains every kept
// voxel whatever the last-bit rounding of its start).
template <int BX, int BY, bool POS>
__device__ __forceinline__ void scatter_box_yf(const TorParams &p, const TileGeom &g,
                                               const TileBox &b, const Event &ev, const Track &t,
                                               bool tof, float scale, int s0, int s1,
                                               double *lds)
{
    constexpr int NC = BX * BY;
    const float eps = 1.0e-4f;
    const float hxr = fmaf(t.hx, p.rfinter_x, eps), dyr = fmaf(t.dymax, p.rfinter_y, eps);
    const float pm = (float)(p.patch - 1);
    const float nxm = (float)(t.nx - 1), nym = (float)(t.ny - 1);
    const float wx0 = (float)b.x0, wx1 = (float)b.x1, wy0 = (float)b.y0, wy1 = (float)b.y1;
    const float sgn = POS ? 1.0f : (scale < 0.0f ? -1.0f : 1.0f);
    const float cthr = ev.thr * ev.cw + __log2f(fabsf(tof ? p.tofk * scale : scale));
    // LDS byte offset of voxel (X, Y) of slice s (y fastest): lsl(s) + 8 (X Wy + Y)
    const int wy8 = 8 * g.Wy;
    const int sstep = 8 * g.Wx * g.Wy;
    int lsl = (s0 - b.sa) * sstep - 8 * (b.x0 * g.Wy + b.y0);
    for (int s = s0; s <= s1; s++, lsl += sstep) {
        Cross c;
        cross_at(p, ev, t, s, c);
        const double sd = (double)s;
        const float cexp = tof ? cthr + (float)fma(fma(t.tc2, sd, t.tc1), sd, t.tc0) : cthr;
        const float fx0 = (float)c.ix0, fy0 = (float)c.iy0;
        const float ib0 = ceilf(fmaf(c.dx0, p.rfinter_x, -hxr));
        const float jb0 = ceilf(fmaf(c.dy0, p.rfinter_y, -dyr));
        const float ilo = fmaxf(fmaxf(ib0, wx0 - fx0), 0.0f);
        const float ihi = fminf(fminf(ib0 + nxm, wx1 - fx0), pm);
        const float jlo = fmaxf(fmaxf(jb0, wy0 - fy0), 0.0f);
        const float jhi = fminf(fminf(jb0 + nym, wy1 - fy0), pm);
        const float ni = ihi - ilo, nj = jhi - jlo;
        float cx2[BX], cux[BX], rt2[BY], ruy[BY];
        {
            float dx = fmaf(-ilo, p.finter_x, c.dx0);
#pragma unroll
            for (int k = 0; k < BX; k++) {
                col_terms(dx, ev.ux, cx2[k], cux[k]);
                cx2[k] = ni >= (float)k ? cx2[k] : INFINITY;
                dx -= p.finter_x;
            }
            float dy = fmaf(-jlo, p.finter_y, c.dy0);
#pragma unroll
            for (int k = 0; k < BY; k++) {
                row_terms(dy, ev.uy, ev.thr, rt2[k], ruy[k]);
                rt2[k] = nj >= (float)k ? rt2[k] : INFINITY;
                dy -= p.finter_y;
            }
        }
        float dt[NC];
        float gap = INFINITY;
#pragma unroll
        for (int jj = 0; jj < BY; jj++) {
#pragma unroll
            for (int ii = 0; ii < BX; ii++) {
                const int k = jj * BX + ii;
                dt[k] = tube_dt(cx2[ii], cux[ii], rt2[jj], ruy[jj]);
                gap = fminf(gap, fabsf(dt[k]));
            }
        }
        const int ib = (int)ilo, jb = (int)jlo;
        // LDS address of the box's first voxel; a clipped-away candidate is
        // never added, so its address may lie outside the tile
        const int a0 = opaque(lsl + __mul24(c.ix0 + ib, wy8) + 8 * (c.iy0 + jb));
        char *const base = reinterpret_cast<char *>(lds);
        if (gap <= kThrMargin) {
            // rare: the reference's FP64 decision, one voxel at a time, for the
            // candidates within the margin of the cut-off; a kept one is added
            // here (the common pass below only adds dt < -margin)
            uint32_t near = 0;
#pragma unroll
            for (int k = 0; k < NC; k++)
                near |= (uint32_t)(fabsf(dt[k]) <= kThrMargin) << k;
            for (uint32_t m = (uint32_t)opaque((int)near); m; m &= m - 1) {
                const int k = __builtin_ctz(m);
                const int i = k % BX, j = k / BX;
                if (inside_exact(p, ev, s, c.ix0 + ib + i, c.iy0 + jb + j)) {
                    const float w = sgn * __builtin_amdgcn_exp2f(fmaf(dt[k], ev.cw, cexp));
                    atomicAdd(reinterpret_cast<double *>(base + a0 + i * wy8 + 8 * j), (double)w);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NC; k++) {
            if (dt[k] < -kThrMargin) {
                const float w = sgn * __builtin_amdgcn_exp2f(fmaf(dt[k], ev.cw, cexp));
                atomicAdd(reinterpret_cast<double *>(base + a0 + (k % BX) * wy8 + 8 * (k / BX)),
                          (double)w);
            }
        }
    }
}

template <int KIND, int NB, bool YF, typename T>
__device__ __forceinline__ float tile_event(const TorParams &p, const TileGeom &g,
                                            const TileBox &b, const Event &ev, const Track &t,
                                            bool tof, float scale, int s0, int s1, T *lds)
{
    constexpr int NBB = NB > 1 ? NB : 2;
    const int shape = box_shape<NB>(t);
    const int P = p.patch;
    const float ts = p.tofk * scale;
    float acc = 0.0f;
    for (int s = s0; s <= s1; s++) {
        Cross c;
        cross_at(p, ev, t, s, c);
        // scatters: the scale rides in amp (one product per slice, not per
        // voxel), with the TOF constant folded into it once
        float amp = KIND == 1 ? (tof ? tof_exp(p, ev, t, s) * ts : scale)
                              : (tof ? tof_factor(p, ev, t, s) : 1.0f);
        if (NB > 0 && shape) {
            PatchIdx idx;
            idx.base = lds_index<YF>(g, b, s, c.ix0, c.iy0);
            idx.si = YF ? g.Wy : 1;
            idx.sj = YF ? 1 : g.Wx;
            const float *src = nullptr;
            if constexpr (KIND == 0)
                src = lds;
            const int il = max(0, b.x0 - c.ix0), ih = min(P - 1, b.x1 - c.ix0);
            const int jl = max(0, b.y0 - c.iy0), jh = min(P - 1, b.y1 - c.iy0);
            constexpr int L = YF ? 2 : 1;
            if (shape == 1)
                acc += slice_box<KIND, NBB, NBB - 1, L>(p, ev, t, c, amp, s, scale, il, ih, jl,
                                                        jh, idx, src, lds);
            else
                acc += slice_box<KIND, NBB, NBB, L>(p, ev, t, c, amp, s, scale, il, ih, jl, jh,
                                                    idx, src, lds);
        } else {
            acc += window_walk<KIND, YF>(p, g, b, ev, t, c, amp, s, scale, lds);
        }
    }
    return acc;
}

// The projection of a plan's events from their walk records (the first
// kernel of a plan's OSEM step): LPE lanes per event as tor_kernel<kProj>
// runs them (the same walk, so the same bits), with the event set-up
// unpacked from the record head instead of formed again from the LOR row by
// each of the event's lanes.  The head's value slot receives the scale of
// the event's backprojection: 1/p for a projection p > 1e-7, else 0 (the
// event adds nothing, as the reference's ratio skips it), so that each of the
// event's listings reads it instead of dividing again.  With pout the
// projection p itself goes to pout[e] instead (the plan's projection op).
template <int NB, int LPE, bool FRAME = false>
__device__ __forceinline__ void record_project(const TorParams &p, float4 *rec, int64_t e,
                                               int sub, bool valid, float *pout = nullptr)
{
    float part = 0.0f;
    Event ev;
    Track t;
    bool any = false;
    if (valid) {
        unpack_walk(p, rec, e, ev, t);
        any = ev.s_lo <= ev.s_hi;
    }
    // The lanes of a wave walk the same slices at the same time (from the
    // wave's first crossed slice, lane sub of an event on slices base + sub
    // + k LPE), so that the gathers of neighbouring events (the plan's event
    // order is a locality order) hit the same cache lines.
    int base = any ? ev.s_lo : 0x7fffffff;
#pragma unroll
    for (int m = 1; m < 64; m <<= 1)
        base = min(base, __shfl_xor(base, m, 64));
    if (any)
        part = walk_event<0, NB, FRAME>(p, ev, t, p.use_tof, 0.0f, base + sub, ev.s_hi, LPE,
                                        ev.s_lo);
    const float pv = group_sum<LPE>(part);
    if (valid && sub == 0) {
        if (pout)
            pout[e] = pv;
        else
            reinterpret_cast<float *>(rec)[4 * kWalkHead * e + kWalkValue] =
                pv > 1e-7 ? 1.0f / pv : 0.0f;
    }
}

// FRAME: p.img is a contiguous op frame of p.img_bytes (frame gathers).
// Workgroup b runs on XCD b % 8 (the dispatcher deals workgroups out round
// robin over the chip's 8 XCDs); remap so that each XCD takes a contiguous
// eighth of the workgroups (of the events, in their locality order), and its
// L2 serves a compact part of the image.  A bijection of [0, gridDim.x).
__device__ __forceinline__ unsigned xcd_block()
{
    const unsigned b = blockIdx.x, n = gridDim.x;
    const unsigned q = n / 8u, r = n % 8u, x = b % 8u;
    return x * q + min(x, r) + b / 8u;
}

// Five waves per SIMD (96 VGPRs) for the 3 x 3 box over a frame at one or
// four lanes per event: the gathers' latency is covered better than at four
// waves (0.38 against 0.41 ms per 1M events); the other forms would spill.
template <int NB, int LPE, bool FRAME>
__global__ __launch_bounds__(256, NB == 3 && FRAME && LPE <= 4 ? 5 : 4) void record_proj_kernel(
    TorParams p, float4 *rec, float *pout)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t e = gid / LPE;
    record_project<NB, LPE, FRAME>(p, rec, e, (int)(gid % LPE), e < p.n, pout);
}


// Scatter of one chunk (modes kBack, kMap, kAdj) into p.out, through the
// FP64 LDS tile ldsd (S * Wy * Wx doubles).  Every thread of the workgroup
// calls it; it synchronises the workgroup before and after the flush.
// VREC (kBack only): the backprojection scales 1/p that the record projection
// stored in the event records (the OSEM step of a plan), not p.val.
// ONE: every listing's event walks the NB x (NB - 1) box (a plan whose
// events all have box shape 1); the other walks are left out of the kernel.
template <int MODE, int NB, bool VREC = false, bool YF = false, bool ONE = false>
__device__ __forceinline__ void scatter_chunk(const TorParams &p, const TileGeom &g,
                                              const TileChunk &ch,
                                              const int *__restrict__ list,
                                              const float4 *__restrict__ rec, double *ldsd)
{
    const TileBox b = tile_box(p, g, ch.tile);
    const int nv = g.S * g.Wy * g.Wx;
    for (int i = threadIdx.x; i < nv; i += blockDim.x)
        ldsd[i] = 0.0;
    __syncthreads();
    const bool tof = MODE != kMap && p.use_tof;
    // The 64 lanes of a wave take listings `stride` apart in the chunk:
    // neighbours in the list are neighbouring (often near-identical) events,
    // whose LDS atomics would hit the same voxels in the same instruction.
    const int stride = (ch.count + 63) >> 6;
    for (int q = threadIdx.x; q < 64 * stride; q += blockDim.x) {
        // sorted plans: consecutive listings (similar slice counts) per wave
        const int j = g.sorted ? q : (q & 63) * stride + (q >> 6);
        if (j >= ch.count)
            continue;
        const int ent = list[ch.start + j];
        const int64_t e = g.spans ? (ent & ((1 << kSpanEventBits) - 1)) : ent;
        Event ev;
        Track t;
        const float rv = unpack_walk(p, rec, e, ev, t);
        int s0, s1;
        if (g.spans) {
            s0 = b.sa + ((ent >> kSpanEventBits) & 15);
            s1 = b.sa + (int)((unsigned)ent >> (kSpanEventBits + 4));
        } else {
            listing_slices(p, b, ev, t, s0, s1);
        }
        const float val = VREC ? rv : p.val[e];
        float scale;
        bool go = true;
        static_assert(!VREC || MODE == kBack, "records hold backprojection scales");
        if (VREC) {
            scale = rv;
            go = rv != 0.0f;
        } else if (MODE == kBack) {
            go = val > 1e-7;
            scale = 1.0f / val;
        } else if (MODE == kMap) {
            scale = 1.0f / val;
        } else {
            scale = val;
            go = scale != 0.0f;
        }
        if (go) {
            constexpr int NBB = NB > 1 ? NB : 2;
            constexpr bool POS = MODE == kBack; // scale 1/v with v > 1e-7
            const int shape = box_shape<NB>(t);
            static_assert(!ONE || (YF && NB > 0), "the single walk is the y-fastest box walk");
            if (ONE || (YF && NB > 0 && shape == 1))
                scatter_box_yf<NBB, NBB - 1, POS>(p, g, b, ev, t, tof, scale, s0, s1, ldsd);
            else if (YF && NB > 0 && shape == 2)
                scatter_box_yf<NBB, NBB, POS>(p, g, b, ev, t, tof, scale, s0, s1, ldsd);
            else
                tile_event<1, NB, YF>(p, g, b, ev, t, tof, scale, s0, s1, ldsd);
        }
    }
    __syncthreads();
    // Flush.  Tile element i = (c * nb + m) * na + f (f along the fastest
    // axis: y for YF, else x; m the middle one; c the slice) and its output
    // offset advance by the block size each step, with carries instead of
    // divisions by the tile extents.
    const int na = YF ? g.Wy : g.Wx, nb = YF ? g.Wx : g.Wy;
    const int oa = YF ? p.oy : p.ox, ob = YF ? p.ox : p.oy;
    const int la = YF ? b.y1 - b.y0 : b.x1 - b.x0, lb = YF ? b.x1 - b.x0 : b.y1 - b.y0;
    const int lc = b.sb - b.sa;
    int f = threadIdx.x % na, m, c;
    {
        const int r = threadIdx.x / na;
        m = r % nb;
        c = r / nb;
    }
    const int df = blockDim.x % na, dr = blockDim.x / na;
    const int dm = dr % nb, dc = dr / nb;
    float *const base = p.out + ((int64_t)b.x0 * p.ox + (int64_t)b.y0 * p.oy + (int64_t)b.sa * p.oz);
    int off = f * oa + m * ob + c * p.oz;
    const int step = df * oa + dm * ob + dc * p.oz;
    const int wrap_f = ob - na * oa, wrap_m = p.oz - nb * ob;
    for (int i = threadIdx.x; i < nv; i += blockDim.x) {
        const float v = (float)ldsd[i];
        if (v != 0.0f && f <= la && m <= lb && c <= lc)
            unsafeAtomicAdd(base + off, v);
        f += df;
        off += step;
        const bool cf = f >= na;
        f -= cf ? na : 0;
        off += cf ? wrap_f : 0;
        m += dm + (cf ? 1 : 0);
        const bool cm = m >= nb;
        m -= cm ? nb : 0;
        off += cm ? wrap_m : 0;
        c += dc + (cm ? 1 : 0);
    }
    __syncthreads();
}

// The lanes of the wave whose v (0 <= v < 64) equals this lane's: one ballot
// per bit of v.  Call with every lane of the wave active.
__device__ __forceinline__ uint64_t lanes_with(int v)
{
    uint64_t m = ~0ull;
#pragma unroll
    for (int b = 0; b < 6; b++) {
        const bool bit = (v >> b) & 1;
        const uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

// Plan build, after the fill pass: order each chunk's listings by the number
// of slices they walk in the tile, longest first (a counting sort in LDS), so
// that the consecutive listings a wave of the scatter takes walk about as
// many slices each.  Inside a bucket the listings keep their order within a
// wave and the waves their order of arrival, which with coalesced reads
// keeps roughly the fill order.  The lanes of a wave with the same key count
// and take their positions together: one LDS atomic per key and wave, the
// lowest such lane's, instead of one per lane on a few shared counters.
// With spans the slice counts come from the entries (the fill pass wrote
// them) instead of the walk records.  One workgroup per chunk, 256 threads;
// LDS: 33 counters and `chunk` ints.  Needs ch.count <= 16 * blockDim.x.
template <bool SPANS>
__device__ __forceinline__ void sort_chunk(const TorParams &p, const TileGeom &g,
                                           const TileChunk &ch, int *__restrict__ list,
                                           const float4 *__restrict__ rec, int *sh)
{
    constexpr int kBuckets = 32, kPer = 16;
    int *cnt = sh, *ord = sh + kBuckets + 1;
    const int lane = (int)threadIdx.x & 63;
    for (int i = threadIdx.x; i <= kBuckets; i += blockDim.x)
        cnt[i] = 0;
    __syncthreads();
    // per entry: its key (-1: no entry), event, and its rank among the wave's
    // lanes with the same key | the first such lane << 6 | their number << 12
    int key[kPer], ev[kPer], grp[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        key[r] = -1;
        ev[r] = 0;
        grp[r] = 0;
        if (r * (int)blockDim.x >= ch.count)
            continue; // the same for every lane of the workgroup
        const int j = r * (int)blockDim.x + (int)threadIdx.x; // coalesced reads
        if (j < ch.count) {
            ev[r] = list[ch.start + j];
            int s0, s1;
            if (SPANS) {
                // the fill pass put the listing's slices in the entry
                s0 = (ev[r] >> kSpanEventBits) & 15;
                s1 = (int)((unsigned)ev[r] >> (kSpanEventBits + 4));
            } else {
                Event e;
                Track t;
                unpack_walk(p, rec, ev[r], e, t);
                listing_slices(p, tile_box(p, g, ch.tile), e, t, s0, s1);
            }
            // longest first: bucket 0 holds the longest walks
            key[r] = kBuckets - 1 - min(max(s1 - s0 + 1, 0), kBuckets - 1);
        }
        const uint64_t m = lanes_with(key[r] + 1);
        const int rank = (int)__builtin_amdgcn_mbcnt_hi(
            (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        const int n = __popcll(m);
        grp[r] = rank | (__builtin_ctzll(m) << 6) | (n << 12);
        if (key[r] >= 0 && rank == 0)
            atomicAdd(&cnt[key[r]], n);
    }
    __syncthreads();
    // exclusive prefix of the bucket counts, by the first wave
    if (threadIdx.x < 64) {
        const int c = lane < kBuckets ? cnt[lane] : 0;
        int v = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(v, d, 64);
            v += lane >= d ? u : 0;
        }
        if (lane < kBuckets)
            cnt[lane] = v - c;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        if (r * (int)blockDim.x >= ch.count)
            continue;
        const int rank = grp[r] & 63;
        int base = 0;
        if (key[r] >= 0 && rank == 0)
            base = atomicAdd(&cnt[key[r]], grp[r] >> 12);
        base = __shfl(base, (grp[r] >> 6) & 63, 64);
        if (key[r] >= 0)
            ord[base + rank] = ev[r];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < ch.count; j += blockDim.x)
        list[ch.start + j] = ord[j];
}

__global__ __launch_bounds__(256) void sort_chunks(TorParams p, TileGeom g,
                                                   const TileChunk *__restrict__ chunks,
                                                   int *__restrict__ list,
                                                   const float4 *__restrict__ rec)
{
    extern __shared__ int sh[];
    if (g.spans)
        sort_chunk<true>(p, g, chunks[blockIdx.x], list, rec, sh);
    else
        sort_chunk<false>(p, g, chunks[blockIdx.x], list, rec, sh);
}

// The chunk sort of many plans with spans in one launch (a plan set build):
// job j sorts the chunks of one plan, workgroups blk0 .. of the grid.
struct SortJob {
    const TileChunk *chunks;
    int *list;
    int64_t blk0;
};

__global__ __launch_bounds__(256) void sort_chunks_many(const SortJob *__restrict__ jobs,
                                                        int njobs)
{
    extern __shared__ int sh[];
    // the plan of this workgroup: the last job whose first workgroup <= b
    const int64_t b = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jobs[mid].blk0 <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const SortJob jb = jobs[lo];
    const TileChunk ch = jb.chunks[b - jb.blk0];
    sort_chunk<true>(TorParams{}, TileGeom{}, ch, jb.list, nullptr, sh);
}

// One workgroup per chunk.  The LDS tile is FP64: LDS FP64 atomic adds run
// at several lanes per cycle where FP32 ones are serialised lane by lane,
// and the tile sums gain precision.  Four waves per SIMD where the walk fits
// in 128 VGPRs without spilling (the LDS atomics need the latency cover);
// the 4 x 4 box walk needs more registers and runs three.  ONE (only the
// NB x (NB - 1) box walk, NB = 3): five waves per SIMD in 96 VGPRs, for
// tiles of at most a fifth of a CU's LDS (0.39 against 0.42 ms per 1M
// events).
template <int MODE, int NB, bool VREC, bool YF, bool ONE = false>
__global__ __launch_bounds__(256, NB == 4 ? (ONE ? 4 : 3) : (ONE ? 5 : 4)) void tile_scatter(
    TorParams p, TileGeom g, const TileChunk *__restrict__ chunks, const int *__restrict__ list,
    const float4 *__restrict__ rec)
{
    extern __shared__ double ldsd[];
    scatter_chunk<MODE, NB, VREC, YF, ONE>(p, g, chunks[blockIdx.x], list, rec, ldsd);
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
        // plans with slot tables keep plain event indices (g.spans 0)
        const int ent = list[ch.start + j];
        const int64_t e = g.spans ? (ent & ((1 << kSpanEventBits) - 1)) : ent;
        Event ev;
        Track t;
        unpack_walk(p, rec, e, ev, t);
        int s0, s1;
        listing_slices(p, b, ev, t, s0, s1);
        const float acc = tile_event<0, NB, false>(p, g, b, ev, t, tof, 0.0f, s0, s1, lds);
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
// The OSEM subset plans of several LOR arrays at once: for request r, the
// plans srfx_tor_plan_create_subsets builds for (lors, n, ld, grid, tile,
// nplans, bounds), written in request order to plans (sum of nplans
// entries).  All requests share the build's phases (one count pass per
// request, one host synchronisation, one fill pass per request, one host
// synchronisation, one chunk sort launch) and two device blocks.
struct SubsetRequest {
    const float *lors;
    int64_t n, ld;
    srfx_grid grid;
    srfx_tile tile;
    int32_t nplans;
    const int64_t *bounds;
};
int plan_create_subsets_many(const SubsetRequest *req, int nreq, int32_t ncols,
                             const srfx_tor_model &m, srfx_tor_plan *plans, hipStream_t st);

} // namespace srfx

// Plan object of srfx_tor_plan_create.
struct srfx_tor_plan_s {
    srfx::TorParams p; // geometry, model and events; image / value pointers set per call
    srfx_grid grid;
    srfx::TileGeom g;
    int cols7;
    int64_t nlist;
    int nchunks;
    int block;
    int *list;             // event of each slot, tiles in order
    float4 *rec;           // walk records (kWalkF4 float4 per event: heads, then tails)
    srfx::TileChunk *chunks;
    float *proj;           // projection buffer of the OSEM step
    int *ev_off, *ev_slot; // tiled projection only: each event's slots
    float *partial;        // tiled projection only: per-slot partial sums
    // Plans built together (srfx_tor_plan_create_subsets) share two device
    // blocks holding their records, projection buffers, lists and chunks;
    // the last plan destroyed frees them.  Null for a plan built alone.
    struct Arena {
        void *mem[2];
        std::atomic<int> refs;
    } *arena;
    std::vector<srfx::TileChunk> hchunks; // host copy of the chunk table
    int one;               // every event walks the NB x (NB - 1) box (box shape 1)
};
