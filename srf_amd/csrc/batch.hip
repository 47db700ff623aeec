// Batched TOF OSEM: many independent reconstructions (different image sizes
// and event counts) advance one subset together.
//
// At creation every (reconstruction, axis group, subset) gets a tile plan of
// its events (tile_core.h).  A subset step is four kernels on one stream:
//   batch_zero     zero every reconstruction's backprojection buffer and the
//                  work counters (two per box class);
//   batch_frames   copy every reconstruction's image into the op frame of
//                  each of its axis groups (frame_core.h), which the
//                  projection reads (a box row of voxels shares a line);
//   batch_project  persistent projection: each block takes the next chunk of
//                  <= 256 events (of any reconstruction and group) from a
//                  counter, walks them from their plan's walk records and
//                  writes their projections into the records' value slots;
//   batch_scatter  persistent tiled backprojection of 1/p: each block takes
//                  the next tile chunk (of any plan), accumulates it in an
//                  FP64 LDS tile and flushes it into that reconstruction's
//                  backprojection buffer;
//   batch_update   x <- x * E * bp for every reconstruction.
// Every block of a persistent kernel exits once its counter passes the end of
// the table.  The step allocates nothing and issues no memset, so a caller can
// capture it in a HIP graph.  Five kernels in all.
#include <algorithm>
#include <new>
#include <vector>

#include "frame_core.h"
#include "tile_core.h"

namespace srfx {

// One (reconstruction, axis group, subset): the subset's events and its plan.
struct BatchJob {
    TorParams p; // img: the reconstruction's image, out: its bp, val: proj
    TileGeom g;
    const int *list;
    const float4 *rec;
    float *proj;
};

struct EvChunk {
    int job, start, count, pad;
};

// Box classes of the batch's jobs: 0 = candidate box 3 with every event on
// the 3 x 2 box and a y-fastest flush (the single-walk scatter), 1 = other
// box-3 jobs, 2 = box 4, 3 = chord walk.  Each class has its own work lists
// and persistent kernels, so that a kernel holds only its class's walk (and
// its registers).
constexpr int kBoxClasses = 4;

inline int box_class(const TorParams &p, bool single_walk)
{
    if (p.box == 3)
        return single_walk && flush_y_fastest(p) ? 0 : 1;
    return p.box == 4 ? 2 : 3;
}

struct TileWork {
    int job, tile, start, count;
};

struct VoxRange {
    float *x;
    const float *e;
    float *bp;
    long long n;
    long long off; // first voxel of this reconstruction in the concatenated range
};

namespace {

constexpr int kChunk = 256;

__device__ __forceinline__ int find_range(const VoxRange *r, int nrec, long long i)
{
    int lo = 0, hi = nrec - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (r[mid].off <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void batch_zero(const VoxRange *__restrict__ r, int nrec,
                                                  long long total, unsigned *counters)
{
    if (blockIdx.x == 0 && threadIdx.x < 2 * kBoxClasses)
        counters[threadIdx.x] = 0u;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        int k = find_range(r, nrec, i);
        r[k].bp[i - r[k].off] = 0.0f;
    }
}

// model/recon_step.py:48-49 for every reconstruction.
__global__ __launch_bounds__(256) void batch_update(const VoxRange *__restrict__ r, int nrec,
                                                    long long total)
{
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        int k = find_range(r, nrec, i);
        long long j = i - r[k].off;
        r[k].x[j] = (r[k].x[j] * r[k].e[j]) * r[k].bp[j];
    }
}

// Next work item of a persistent block (the same value in every thread).
__device__ __forceinline__ int next_item(unsigned *counter)
{
    __shared__ int item;
    if (threadIdx.x == 0)
        item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int c = __builtin_amdgcn_readfirstlane(item);
    __syncthreads();
    return c;
}

// Next work item of a persistent wave (the same value in every lane; every
// lane active).
__device__ __forceinline__ int next_wave_item(unsigned *counter)
{
    int item = 0;
    if (__lane_id() == 0)
        item = (int)atomicAdd(counter, 1u);
    return __builtin_amdgcn_readfirstlane(item);
}

// Waves pull work items of T events from the event chunks (kChunk / T items
// per chunk) of one box class (NB), walked in rounds of 64 / LPE events at
// LPE lanes per event, without synchronising with the other waves of their
// block, so a wave with long walks holds up no other.  The default schedule,
// LPE = 4 and T = 64 (quarter chunks), is the projection schedule of the plan
// OSEM step.  The gathers read the jobs' contiguous op frames
// (record_project<.., true>).  Five waves per SIMD for the 3 x 3 box at one
// lane per event and items of up to 128 events (a 12-byte spill outside the
// walk; C4 iteration 88 against 93 ms); the other schedules measured slower
// at five and keep four.
template <int NB, int LPE, int T>
__global__ __launch_bounds__(256, NB == 3 && LPE == 1 && T <= 128 ? 5 : 4) void batch_project(
    const BatchJob *__restrict__ jobs, const EvChunk *__restrict__ chunks, int nchunks,
    unsigned *counter)
{
    constexpr int kParts = kChunk / T, kRound = 64 / LPE;
    static_assert(kChunk % T == 0 && T % kRound == 0, "items split chunks into whole rounds");
    for (;;) {
        const int q = next_wave_item(counter);
        if (q >= kParts * nchunks)
            break;
        const EvChunk ch = chunks[q / kParts];
        const BatchJob &job = jobs[__builtin_amdgcn_readfirstlane(ch.job)];
        float4 *rec = const_cast<float4 *>(job.rec);
        const int lane = (int)__lane_id(), sub = lane % LPE;
        for (int r = 0; r < T; r += kRound) {
            const int i = (q % kParts) * T + r + lane / LPE;
            const bool valid = i < ch.count;
            const int64_t e = (int64_t)ch.start + i;
            record_project<NB, LPE, true>(job.p, rec, e, sub, valid);
        }
    }
}

// Blocks pull tile chunks of one box class (NB); four waves per SIMD where
// the walk fits, five for the single-walk class (ONE), as tile_scatter.
template <int NB, bool ONE = false>
__global__ __launch_bounds__(256, ONE ? 5 : (NB == 4 ? 3 : 4)) void batch_scatter(
    const BatchJob *__restrict__ jobs, const TileWork *__restrict__ work, int nwork,
    unsigned *counter)
{
    extern __shared__ double ldsd[];
    for (;;) {
        const int c = next_item(counter);
        if (c >= nwork)
            break;
        const TileWork w = work[c];
        const BatchJob &job = jobs[__builtin_amdgcn_readfirstlane(w.job)];
        const TileChunk ch{w.tile, w.start, w.count, 0};
        if constexpr (ONE)
            scatter_chunk<kBack, NB, true, true, true>(job.p, job.g, ch, job.list, job.rec, ldsd);
        else if (flush_y_fastest(job.p))
            scatter_chunk<kBack, NB, true, true>(job.p, job.g, ch, job.list, job.rec, ldsd);
        else
            scatter_chunk<kBack, NB, true, false>(job.p, job.g, ch, job.list, job.rec, ldsd);
    }
}

// One op-frame copy: the world image read with the group's op-frame strides
// into a contiguous [gz][gy][gx] buffer, tiles [first, first + tx * ty * gz)
// of the batch_frames grid.
struct FrameJob {
    const float *src;
    float *dst;
    int gx, gy, gz, tx, ty;
    int64_t sx, sy, sz;
    long long first;
};

__global__ __launch_bounds__(256) void batch_frames(const FrameJob *__restrict__ jobs, int njobs)
{
    __shared__ float tile[kFrameTile][kFrameTile + 1];
    const long long blk = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) { // last job whose first tile is <= blk
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= blk)
            lo = mid;
        else
            hi = mid - 1;
    }
    const FrameJob &f = jobs[lo];
    const long long t = blk - f.first;
    const int bx = (int)(t % f.tx);
    const long long r = t / f.tx;
    const int by = (int)(r % f.ty);
    frame_tile(f.src, f.gx, f.gy, f.sx, f.sy, f.sz, f.dst, bx * kFrameTile, by * kFrameTile,
               r / f.ty, tile);
}

// Op frame of axis group a for a C-order (nx, ny, nz) world image, as
// physics/split_lors.py:54-85 derives it by transposing: op (x, y, z) =
// world (y, z, x) for the x group, (x, z, y) for y, (x, y, z) for z.
srfx_grid group_grid(const srfx_recon &r, int a)
{
    static const int perm[3][3] = {{1, 2, 0}, {0, 2, 1}, {0, 1, 2}};
    const int64_t st[3] = {(int64_t)r.shape[1] * r.shape[2], (int64_t)r.shape[2], 1};
    srfx_grid g;
    for (int k = 0; k < 3; k++) {
        int w = perm[a][k];
        g.shape[k] = r.shape[w];
        g.center[k] = r.center[w];
        g.size[k] = r.size[w];
        g.stride[k] = st[w];
    }
    return g;
}

// Tile of the batch's plans: 12 slices x 32 x 10 voxels, <= 2048 events per
// workgroup (the FP64 LDS tile is 30 KB: five workgroups per CU).
constexpr srfx_tile kBatchTile = {12, 32, 10, 2048, 256, 0};

} // namespace
} // namespace srfx

struct srfx_tor_batch {
    int nrec = 0, nsub = 0, njobs = 0;
    std::vector<srfx_tor_plan> plans;
    srfx::BatchJob *d_jobs = nullptr;
    // work lists per (subset, box class): index subset * kBoxClasses + class
    std::vector<srfx::EvChunk *> d_ev;
    std::vector<int> nev;
    std::vector<srfx::TileWork *> d_work;
    std::vector<int> nwork;
    srfx::VoxRange *d_vox = nullptr;
    long long total_vox = 0;
    unsigned *d_counters = nullptr;
    // the block holding the op-frame copies, one per (reconstruction,
    // non-empty group)
    std::vector<float *> frames;
    srfx::FrameJob *d_frames = nullptr;
    int nframes = 0;
    long long frame_blocks = 0;
    unsigned scatter_lds = 0;
    int proj_blocks[srfx::kBoxClasses] = {}, scatter_blocks[srfx::kBoxClasses] = {};
    int elem_blocks = 0;
    int cus = 0;
    int proj_lanes = 4, proj_events = 64; // projection schedule (batch_project<LPE, T>)
};

using namespace srfx;

// The persistent projection of schedule (LPE, T): with subset < 0 sets the
// plan's grid from the kernel's occupancy, else enqueues it for that subset.
template <int NB, int LPE, int T>
static int project_class(srfx_tor_batch *b, int c, int subset, hipStream_t st)
{
    if (subset < 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_project<NB, LPE, T>, 256,
                                                         0) != hipSuccess)
            return SRFX_ELAUNCH;
        b->proj_blocks[c] = std::max(1, b->cus * std::max(1, per_cu));
        return SRFX_OK;
    }
    const int w = subset * kBoxClasses + c;
    if (b->nev[w] > 0)
        hipLaunchKernelGGL((batch_project<NB, LPE, T>), dim3(b->proj_blocks[c]), dim3(256), 0, st,
                           b->d_jobs, b->d_ev[w], b->nev[w], b->d_counters + 2 * c);
    return SRFX_OK;
}

template <int LPE, int T>
static int project_launch(srfx_tor_batch *b, int subset, hipStream_t st)
{
    int rc = project_class<3, LPE, T>(b, 0, subset, st);
    if (!rc)
        rc = project_class<3, LPE, T>(b, 1, subset, st);
    if (!rc)
        rc = project_class<4, LPE, T>(b, 2, subset, st);
    if (!rc)
        rc = project_class<0, LPE, T>(b, 3, subset, st);
    return rc;
}

// Lanes per event {1, 4, 16} x events per work item {64, 128, 256}.
static int project_schedule(srfx_tor_batch *b, int lanes, int events, int subset, hipStream_t st)
{
    if (lanes == 1 && events == 64)
        return project_launch<1, 64>(b, subset, st);
    if (lanes == 1 && events == 128)
        return project_launch<1, 128>(b, subset, st);
    if (lanes == 1 && events == 256)
        return project_launch<1, 256>(b, subset, st);
    if (lanes == 4 && events == 64)
        return project_launch<4, 64>(b, subset, st);
    if (lanes == 4 && events == 128)
        return project_launch<4, 128>(b, subset, st);
    if (lanes == 4 && events == 256)
        return project_launch<4, 256>(b, subset, st);
    if (lanes == 16 && events == 64)
        return project_launch<16, 64>(b, subset, st);
    if (lanes == 16 && events == 128)
        return project_launch<16, 128>(b, subset, st);
    if (lanes == 16 && events == 256)
        return project_launch<16, 256>(b, subset, st);
    return SRFX_EINVAL;
}

static void free_plan(srfx_tor_batch *b)
{
    if (!b)
        return;
    (void)hipDeviceSynchronize();
    for (srfx_tor_plan p : b->plans)
        (void)srfx_tor_plan_destroy(p);
    (void)hipFree(b->d_jobs);
    for (EvChunk *c : b->d_ev)
        (void)hipFree(c);
    for (TileWork *w : b->d_work)
        (void)hipFree(w);
    (void)hipFree(b->d_vox);
    (void)hipFree(b->d_counters);
    for (float *f : b->frames)
        (void)hipFree(f);
    (void)hipFree(b->d_frames);
    delete b;
}

template <class T> static bool upload(T **dst, const std::vector<T> &v)
{
    *dst = nullptr;
    if (v.empty())
        return true;
    return hipMalloc((void **)dst, sizeof(T) * v.size()) == hipSuccess &&
           hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) == hipSuccess;
}

extern "C" int srfx_tor_batch_create(const srfx_recon *recons, int32_t nrecon,
                                     int32_t nb_subsets, srfx_tor_model m,
                                     srfx_tor_batch **out)
{
    if (!out || !recons || nrecon <= 0 || nb_subsets <= 0)
        return SRFX_EINVAL;
    *out = nullptr;
    for (int r = 0; r < nrecon; r++) {
        const srfx_recon &rc = recons[r];
        if (!rc.image || !rc.emap || !rc.bp)
            return SRFX_EINVAL;
        for (int a = 0; a < 3; a++) {
            TorParams p;
            int st = setup(p, rc.lors[a], rc.n[a], rc.ld[a], group_grid(rc, a), m.kernel_width,
                           m.geo_sigma2_flag, m.tof_bin, m.tof_sigma2);
            if (st)
                return st;
            // the frame gathers address an op frame with 32-bit byte offsets
            // and 24-bit slice products
            const srfx_grid gg = group_grid(rc, a);
            if (4 * (int64_t)rc.shape[0] * rc.shape[1] * rc.shape[2] >= ((int64_t)1 << 31) ||
                (int64_t)gg.shape[0] * gg.shape[1] >= ((int64_t)1 << 24))
                return SRFX_EUNSUPPORTED;
        }
    }
    srfx_tor_batch *b = new (std::nothrow) srfx_tor_batch;
    if (!b)
        return SRFX_ENOMEM;
    b->nrec = nrecon;
    b->nsub = nb_subsets;
    std::vector<BatchJob> jobs;
    std::vector<std::vector<EvChunk>> ev((size_t)nb_subsets * kBoxClasses);
    std::vector<std::vector<TileWork>> work((size_t)nb_subsets * kBoxClasses);
    std::vector<VoxRange> vox;
    std::vector<FrameJob> frames;
    long long off = 0;
    int rc_status = SRFX_OK;
    // one block for the op-frame copies of every (reconstruction, non-empty
    // group), carved in the loop below
    size_t frame_floats = 0;
    for (int r = 0; r < nrecon; r++)
        for (int a = 0; a < 3; a++)
            if (recons[r].n[a] > 0)
                frame_floats += (size_t)recons[r].shape[0] * recons[r].shape[1] *
                                recons[r].shape[2];
    float *frame_next = nullptr;
    if (frame_floats > 0) {
        if (hipMalloc((void **)&frame_next, sizeof(float) * frame_floats) != hipSuccess) {
            delete b;
            return SRFX_ENOMEM;
        }
        b->frames.push_back(frame_next);
    }
    // Per subset: contiguous rows [k n / K, (k+1) n / K) of every group
    // (graph/reconstruction/worker.py:79-85), each with its own plan; the
    // plans of every (reconstruction, group) are built in one call.
    std::vector<std::vector<int64_t>> all_bounds((size_t)nrecon * 3,
                                                 std::vector<int64_t>(nb_subsets + 1));
    std::vector<SubsetRequest> requests;
    for (int r = 0; r < nrecon; r++)
        for (int a = 0; a < 3; a++) {
            const srfx_recon &rc = recons[r];
            std::vector<int64_t> &bd = all_bounds[(size_t)r * 3 + a];
            for (int k = 0; k <= nb_subsets; k++)
                bd[k] = k * rc.n[a] / nb_subsets;
            requests.push_back(SubsetRequest{rc.lors[a], rc.n[a], rc.ld[a], group_grid(rc, a),
                                             kBatchTile, nb_subsets, bd.data()});
        }
    std::vector<srfx_tor_plan> all_plans(requests.size() * nb_subsets, nullptr);
    rc_status = plan_create_subsets_many(requests.data(), (int)requests.size(), 10, m,
                                         all_plans.data(), nullptr);
    if (rc_status) {
        free_plan(b);
        return rc_status;
    }
    // the batch owns every plan from here (free_plan), the empty ones too:
    // they hold no device memory of their own, and destroying them here would
    // synchronise the device once each
    b->plans = all_plans;
    for (int r = 0; r < nrecon && !rc_status; r++) {
        const srfx_recon &rc = recons[r];
        for (int a = 0; a < 3 && !rc_status; a++) {
            const int64_t n = rc.n[a];
            // the group's op-frame copy of the image, read by its projections
            float *frame = nullptr;
            const srfx_grid gg = group_grid(rc, a);
            if (n > 0) {
                const size_t nv = (size_t)rc.shape[0] * rc.shape[1] * rc.shape[2];
                frame = frame_next;
                frame_next += nv;
                FrameJob f;
                f.src = rc.image;
                f.dst = frame;
                f.gx = gg.shape[0];
                f.gy = gg.shape[1];
                f.gz = gg.shape[2];
                f.tx = (f.gx + kFrameTile - 1) / kFrameTile;
                f.ty = (f.gy + kFrameTile - 1) / kFrameTile;
                f.sx = gg.stride[0];
                f.sy = gg.stride[1];
                f.sz = gg.stride[2];
                f.first = b->frame_blocks;
                b->frame_blocks += (long long)f.tx * f.ty * f.gz;
                frames.push_back(f);
            }
            // the group's subset plans
            const std::vector<int64_t> &bounds = all_bounds[(size_t)r * 3 + a];
            const srfx_tor_plan *sub = all_plans.data() + ((size_t)r * 3 + a) * nb_subsets;
            for (int k = 0; k < nb_subsets; k++) {
                const int64_t lo = bounds[k], hi = bounds[k + 1];
                srfx_tor_plan plan = sub[k];
                if (hi <= lo)
                    continue;
                BatchJob job;
                job.p = plan->p;
                job.p.img = frame; // contiguous op frame, read by the frame gathers
                job.p.sx = 1;
                job.p.sy = gg.shape[0];
                job.p.sz = gg.shape[0] * gg.shape[1];
                job.p.img_bytes = (int)(4 * (int64_t)job.p.sz * gg.shape[2]);
                const int bc = box_class(job.p, plan->one != 0);
                job.p.out = rc.bp;
                job.p.val = plan->proj;
                job.g = plan->g;
                job.list = plan->list;
                job.rec = plan->rec;
                job.proj = plan->proj;
                const int j = (int)jobs.size();
                jobs.push_back(job);
                for (int64_t s = 0; s < hi - lo; s += kChunk)
                    ev[(size_t)k * kBoxClasses + bc].push_back(
                        EvChunk{j, (int)s, (int)std::min<int64_t>(kChunk, hi - lo - s), 0});
                // the plan's chunk table, kept on the host by its build
                const std::vector<TileChunk> &ch = plan->hchunks;
                if ((int)ch.size() != plan->nchunks) {
                    rc_status = SRFX_ELAUNCH;
                    break;
                }
                for (const TileChunk &c : ch)
                    work[(size_t)k * kBoxClasses + bc].push_back(
                        TileWork{j, c.tile, c.start, c.count});
            }
        }
        long long nv = (long long)rc.shape[0] * rc.shape[1] * rc.shape[2];
        vox.push_back(VoxRange{rc.image, rc.emap, rc.bp, nv, off});
        off += nv;
    }
    if (rc_status) {
        free_plan(b);
        return rc_status;
    }
    b->njobs = (int)jobs.size();
    b->total_vox = off;
    b->nframes = (int)frames.size();
    bool ok = upload(&b->d_jobs, jobs) && upload(&b->d_vox, vox) && upload(&b->d_frames, frames) &&
              hipMalloc(&b->d_counters, 2 * kBoxClasses * sizeof(unsigned)) == hipSuccess;
    for (size_t w = 0; ok && w < ev.size(); w++) {
        EvChunk *e = nullptr;
        TileWork *t = nullptr;
        ok = upload(&e, ev[w]) && upload(&t, work[w]);
        b->d_ev.push_back(e);
        b->nev.push_back((int)ev[w].size());
        b->d_work.push_back(t);
        b->nwork.push_back((int)work[w].size());
    }
    b->scatter_lds = (unsigned)(sizeof(double) * kBatchTile.slices * kBatchTile.wx * kBatchTile.wy);
    int dev = 0, cus = 0, per_cu_s[kBoxClasses] = {};
    ok = ok && hipGetDevice(&dev) == hipSuccess &&
         hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
         hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_s[0], batch_scatter<3, true>, 256,
                                                      b->scatter_lds) == hipSuccess &&
         hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_s[1], batch_scatter<3>, 256,
                                                      b->scatter_lds) == hipSuccess &&
         hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_s[2], batch_scatter<4>, 256,
                                                      b->scatter_lds) == hipSuccess &&
         hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_s[3], batch_scatter<0>, 256,
                                                      b->scatter_lds) == hipSuccess;
    b->cus = cus;
    ok = ok && project_schedule(b, b->proj_lanes, b->proj_events, -1, nullptr) == SRFX_OK;
    if (!ok) {
        free_plan(b);
        return SRFX_ELAUNCH;
    }
    for (int c = 0; c < kBoxClasses; c++)
        b->scatter_blocks[c] = std::max(1, cus * std::max(1, per_cu_s[c]));
    b->elem_blocks = std::max(1, cus * 4);
    *out = b;
    return SRFX_OK;
}

extern "C" int srfx_tor_batch_subset(srfx_tor_batch *b, int32_t subset, void *stream)
{
    if (!b || subset < 0 || subset >= b->nsub)
        return SRFX_EINVAL;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(batch_zero, dim3(b->elem_blocks), dim3(256), 0, st, b->d_vox, b->nrec,
                       b->total_vox, b->d_counters);
    if (b->nframes > 0)
        hipLaunchKernelGGL(batch_frames, dim3((unsigned)b->frame_blocks), dim3(256), 0, st,
                           b->d_frames, b->nframes);
    project_schedule(b, b->proj_lanes, b->proj_events, subset, st);
    for (int c = 0; c < kBoxClasses; c++) {
        const int w = subset * kBoxClasses + c;
        if (b->nwork[w] == 0)
            continue;
        if (c == 0)
            hipLaunchKernelGGL((batch_scatter<3, true>), dim3(b->scatter_blocks[c]), dim3(256),
                               b->scatter_lds, st, b->d_jobs, b->d_work[w], b->nwork[w],
                               b->d_counters + 2 * c + 1);
        else if (c == 1)
            hipLaunchKernelGGL(batch_scatter<3>, dim3(b->scatter_blocks[c]), dim3(256),
                               b->scatter_lds, st, b->d_jobs, b->d_work[w], b->nwork[w],
                               b->d_counters + 2 * c + 1);
        else if (c == 2)
            hipLaunchKernelGGL(batch_scatter<4>, dim3(b->scatter_blocks[c]), dim3(256),
                               b->scatter_lds, st, b->d_jobs, b->d_work[w], b->nwork[w],
                               b->d_counters + 2 * c + 1);
        else
            hipLaunchKernelGGL(batch_scatter<0>, dim3(b->scatter_blocks[c]), dim3(256),
                               b->scatter_lds, st, b->d_jobs, b->d_work[w], b->nwork[w],
                               b->d_counters + 2 * c + 1);
    }
    hipLaunchKernelGGL(batch_update, dim3(b->elem_blocks), dim3(256), 0, st, b->d_vox, b->nrec,
                       b->total_vox);
    return launch_status();
}

extern "C" int srfx_tor_batch_set_schedule(srfx_tor_batch *b, int32_t lanes_per_event,
                                           int32_t events_per_item)
{
    if (!b)
        return SRFX_EINVAL;
    // the grid changes only when the schedule exists and its occupancy query succeeds
    const int rc = project_schedule(b, lanes_per_event, events_per_item, -1, nullptr);
    if (rc == SRFX_OK) {
        b->proj_lanes = lanes_per_event;
        b->proj_events = events_per_item;
    }
    return rc;
}

extern "C" int srfx_tor_batch_destroy(srfx_tor_batch *b)
{
    free_plan(b);
    return SRFX_OK;
}
