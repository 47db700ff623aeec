// Batched TOF OSEM: many independent reconstructions (different image sizes
// and event counts) advance one subset in one persistent launch.
//
// A plan holds, in device memory, one TorParams per (reconstruction, axis
// group) and, per subset, a table of work chunks of at most 256 events that
// never straddle two groups.  A subset step is three kernels on one stream:
//   batch_zero    zero every reconstruction's backprojection buffer and the
//                 work counter;
//   batch_osem    persistent fused projection + ratio backprojection: each
//                 block takes the next chunk from the counter until none is
//                 left (every block exits once the counter passes the table);
//   batch_update  x <- x * E * bp for every reconstruction.
// The step allocates nothing and issues no memset, so a caller can capture it
// in a HIP graph.
#include <algorithm>
#include <new>
#include <vector>

#include "tor_core.h"

namespace srfx {
namespace {

constexpr int kChunk = 256;

struct Chunk {
    int job;
    int start;
    int count;
    int pad;
};

struct VoxRange {
    float *x;
    const float *e;
    float *bp;
    long long n;
    long long off; // first voxel of this reconstruction in the concatenated range
};

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
                                                  long long total, unsigned *counter)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *counter = 0u;
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

__global__ __launch_bounds__(256) void batch_osem(const TorParams *__restrict__ jobs,
                                                  const Chunk *__restrict__ chunks, int nchunks,
                                                  unsigned *counter)
{
    __shared__ int item;
    for (;;) {
        if (threadIdx.x == 0)
            item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int c = __builtin_amdgcn_readfirstlane(item);
        __syncthreads();
        if (c >= nchunks)
            break;
        const Chunk ch = chunks[c];
        const TorParams p = jobs[__builtin_amdgcn_readfirstlane(ch.job)];
        const bool valid = (int)threadIdx.x < ch.count;
        const int64_t e = (int64_t)ch.start + threadIdx.x;
        switch (p.box) {
        case 3:
            run_event<kOsem, 3, 1>(p, e, 0, valid);
            break;
        case 4:
            run_event<kOsem, 4, 1>(p, e, 0, valid);
            break;
        default:
            run_event<kOsem, 0, 1>(p, e, 0, valid);
            break;
        }
    }
}

} // namespace
} // namespace srfx

struct srfx_tor_batch {
    int nrec = 0, nsub = 0, njobs = 0;
    srfx::TorParams *d_jobs = nullptr;
    std::vector<srfx::Chunk *> d_chunks;
    std::vector<int> nchunks;
    srfx::VoxRange *d_vox = nullptr;
    long long total_vox = 0;
    unsigned *d_counter = nullptr;
    int osem_blocks = 0, elem_blocks = 0;
};

using namespace srfx;

static void free_plan(srfx_tor_batch *b)
{
    if (!b)
        return;
    (void)hipFree(b->d_jobs);
    for (Chunk *c : b->d_chunks)
        (void)hipFree(c);
    (void)hipFree(b->d_vox);
    (void)hipFree(b->d_counter);
    delete b;
}

// Op frame of axis group a for a C-order (nx, ny, nz) world image, as
// physics/split_lors.py:54-85 derives it by transposing: op (x, y, z) =
// world (y, z, x) for the x group, (x, z, y) for y, (x, y, z) for z.
static srfx_grid group_grid(const srfx_recon &r, int a)
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

extern "C" int srfx_tor_batch_create(const srfx_recon *recons, int32_t nrecon,
                                     int32_t nb_subsets, srfx_tor_model m,
                                     srfx_tor_batch **out)
{
    if (!out || !recons || nrecon <= 0 || nb_subsets <= 0)
        return SRFX_EINVAL;
    *out = nullptr;
    std::vector<TorParams> jobs;
    std::vector<VoxRange> vox;
    long long off = 0;
    for (int r = 0; r < nrecon; r++) {
        const srfx_recon &rc = recons[r];
        if (!rc.image || !rc.emap || !rc.bp)
            return SRFX_EINVAL;
        for (int a = 0; a < 3; a++) {
            TorParams p;
            srfx_grid g = group_grid(rc, a);
            int st = setup(p, rc.lors[a], rc.n[a], rc.ld[a], g, m.kernel_width,
                           m.geo_sigma2_flag, m.tof_bin, m.tof_sigma2);
            if (st)
                return st;
            p.img = rc.image;
            p.out = rc.bp;
            jobs.push_back(p);
        }
        long long nv = (long long)rc.shape[0] * rc.shape[1] * rc.shape[2];
        vox.push_back(VoxRange{rc.image, rc.emap, rc.bp, nv, off});
        off += nv;
    }
    srfx_tor_batch *b = new (std::nothrow) srfx_tor_batch;
    if (!b)
        return SRFX_ELAUNCH;
    b->nrec = nrecon;
    b->nsub = nb_subsets;
    b->njobs = (int)jobs.size();
    b->total_vox = off;
    bool ok = hipMalloc(&b->d_jobs, sizeof(TorParams) * jobs.size()) == hipSuccess &&
              hipMalloc(&b->d_vox, sizeof(VoxRange) * vox.size()) == hipSuccess &&
              hipMalloc(&b->d_counter, sizeof(unsigned)) == hipSuccess &&
              hipMemcpy(b->d_jobs, jobs.data(), sizeof(TorParams) * jobs.size(),
                        hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(b->d_vox, vox.data(), sizeof(VoxRange) * vox.size(),
                        hipMemcpyHostToDevice) == hipSuccess;
    // Per subset: contiguous rows [k n / K, (k+1) n / K) of every group
    // (graph/reconstruction/worker.py:79-85), cut into chunks.
    for (int k = 0; ok && k < nb_subsets; k++) {
        std::vector<Chunk> ch;
        for (int j = 0; j < b->njobs; j++) {
            int64_t n = jobs[j].n;
            int64_t lo = k * n / nb_subsets, hi = (k + 1) * n / nb_subsets;
            for (int64_t s = lo; s < hi; s += kChunk)
                ch.push_back(Chunk{j, (int)s, (int)std::min<int64_t>(kChunk, hi - s), 0});
        }
        Chunk *d = nullptr;
        if (!ch.empty()) {
            ok = hipMalloc(&d, sizeof(Chunk) * ch.size()) == hipSuccess &&
                 hipMemcpy(d, ch.data(), sizeof(Chunk) * ch.size(), hipMemcpyHostToDevice) ==
                     hipSuccess;
        }
        b->d_chunks.push_back(d);
        b->nchunks.push_back((int)ch.size());
    }
    int dev = 0, cus = 0, per_cu = 0;
    ok = ok && hipGetDevice(&dev) == hipSuccess &&
         hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
         hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_osem, 256, 0) ==
             hipSuccess;
    if (!ok) {
        free_plan(b);
        return SRFX_ELAUNCH;
    }
    b->osem_blocks = std::max(1, cus * std::max(1, per_cu));
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
                       b->total_vox, b->d_counter);
    if (b->nchunks[subset] > 0)
        hipLaunchKernelGGL(batch_osem, dim3(b->osem_blocks), dim3(256), 0, st, b->d_jobs,
                           b->d_chunks[subset], b->nchunks[subset], b->d_counter);
    hipLaunchKernelGGL(batch_update, dim3(b->elem_blocks), dim3(256), 0, st, b->d_vox, b->nrec,
                       b->total_vox);
    return launch_status();
}

extern "C" int srfx_tor_batch_destroy(srfx_tor_batch *b)
{
    free_plan(b);
    return SRFX_OK;
}
