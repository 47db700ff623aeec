/*
 * srfx.h — C ABI of the MI355X-native SRF projector library (libsrfx.so).
 *
 * Every entry point takes device pointers, plain sizes and a HIP stream
 * (passed as void*, NULL = the default stream), enqueues its work on that
 * stream without synchronising, and returns an int status (SRFX_OK = 0).
 * Geometry is passed by value; nothing is copied back to the host.  The
 * library holds no state between calls.
 *
 * Reference counterparts are TensorFlow custom ops of Hong-Xiang/SRF; paths
 * below are relative to the reference repository root.
 *
 * Frames and layouts
 *   LOR arrays are structure-of-arrays: column c of event e is lors[c*ld + e]
 *   (the [ncols, N] tensor the reference ops slice row by row, tor.cc:140-160,
 *   siddon.cc:108-122).  ld >= n.
 *   An op works in "its" frame: x, y are the in-slice axes and z the slice axis
 *   for TOR (the reference passes grid/center/size reversed and permutes the
 *   image per axis group, physics/split_lors.py:54-89).  srfx_grid.stride gives
 *   the element stride of one step along the op's x, y and z; all-zero strides
 *   mean the reference layout, a contiguous [gz][gy][gx] array.  Non-zero
 *   strides let a caller index an image stored in another axis order (for
 *   example the unpermuted (nx, ny, nz) image) without transposing it.
 *
 * Outputs
 *   By default an op first zeroes its output, as the reference ops do with
 *   cudaMemset (tor.cc:132, 222, 310; siddon.cc:106, 171, 242).  With
 *   SRFX_ACCUMULATE in flags the op adds into the output instead, which is how
 *   the three TOR axis groups are summed into one image.  SRFX_FULL_PATCH
 *   selects the reference's full patch loop in the TOR ops.
 *
 * Floating point
 *   Scatter ops use FP32 atomics: the order of additions, and so the last bits
 *   of the result, vary from run to run, as in the reference.
 */
#ifndef SRFX_H
#define SRFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum srfx_status {
    SRFX_OK = 0,
    SRFX_EINVAL = 1,      /* bad argument: null pointer, n < 0, ld < n, grid <= 0, size <= 0 */
    SRFX_ELAUNCH = 2,     /* a HIP call or kernel launch failed */
    SRFX_EUNSUPPORTED = 3, /* valid but unsupported (e.g. TOR patch wider than 16 voxels) */
    SRFX_ENOMEM = 4,       /* device workspace allocation failed */
    SRFX_ETOOLARGE = 5     /* input larger than a no-allocation entry point's static scratch */
};

#define SRFX_ACCUMULATE 1u
/* TOR ops: evaluate every voxel of the reference's P x P patch (its loop
 * structure) instead of only the voxels on the tube's chords.  Same results;
 * slower; kept for verification. */
#define SRFX_FULL_PATCH 2u
/* TOR ops: launch schedule.  Lanes that share one event (its crossed slices
 * interleaved between them): SRFX_LPE_1, SRFX_LPE_4, SRFX_LPE_16; without a
 * flag one lane per event, except the projection of a plan's OSEM step
 * (srfx_tor_plan_osem_subset), which takes four.
 * Threads per block: SRFX_BLOCK_256 (default), SRFX_BLOCK_64, SRFX_BLOCK_128.
 * Results agree to FP32 rounding across schedules. */
#define SRFX_LPE_SHIFT 4
#define SRFX_LPE_1 (1u << SRFX_LPE_SHIFT)
#define SRFX_LPE_4 (2u << SRFX_LPE_SHIFT)
#define SRFX_LPE_16 (3u << SRFX_LPE_SHIFT)
#define SRFX_BLOCK_SHIFT 6
#define SRFX_BLOCK_64 (1u << SRFX_BLOCK_SHIFT)
#define SRFX_BLOCK_128 (2u << SRFX_BLOCK_SHIFT)
#define SRFX_BLOCK_256 (3u << SRFX_BLOCK_SHIFT)
/* Slab schedule of the TOR ops (no counterpart in the reference, which
 * launches one kernel per slice, tor.cu.cc:377-391): the slices are cut into
 * contiguous slabs and each event is walked once per slab, the slab's part
 * only.  SRFX_SLAB_XCD8 puts workgroup b on slab b % 8, so that each of the
 * chip's 8 XCDs gathers from and scatters into one eighth of the image;
 * SRFX_SLAB_BLOCKED8 is the same work without that affinity; SRFX_SLAB_XCD16
 * uses 16 slabs, two per XCD.  Projection and OSEM then take a stream-ordered
 * workspace of (slabs x n) floats from the device's memory pool.  Lanes per
 * event is 1 under a slab schedule. */
#define SRFX_SLAB_SHIFT 8
#define SRFX_SLAB_XCD8 (1u << SRFX_SLAB_SHIFT)
#define SRFX_SLAB_BLOCKED8 (2u << SRFX_SLAB_SHIFT)
#define SRFX_SLAB_XCD16 (3u << SRFX_SLAB_SHIFT)
/* srfx_tor_plan_osem_subset: compute the projection half with the tiled
 * gather instead of the untiled kernel (selected by the schedule bits). */
#define SRFX_TILED_PROJECTION (1u << 10)
/* srfx_tor_plan_osem_subset in two calls (e.g. to time the two kernels):
 * SRFX_OSEM_PROJECT_ONLY runs only the projection, which stores each
 * event's backprojection scale 1/p (0 for p <= 1e-7) in the plan's event
 * records, SRFX_OSEM_SCATTER_ONLY only the tiled backprojection of 1/p from
 * them (zeroing bp unless SRFX_ACCUMULATE).  Not with SRFX_TILED_PROJECTION. */
#define SRFX_OSEM_PROJECT_ONLY (1u << 11)
#define SRFX_OSEM_SCATTER_ONLY (1u << 12)
/* srfx_tor_plan_osem_subset: the image argument is the op-frame copy of the
 * image, contiguous [gz][gy][gx] (what the reference's transpose into the
 * group frame produces, split_lors.py:79), whatever strides the plan was
 * built with.  The box walk then reads each voxel row from one or two cache
 * lines instead of one line per voxel column. */
#define SRFX_IMAGE_OP_FRAME (1u << 13)

typedef struct srfx_grid {
    int32_t shape[3];  /* gx, gy, gz: voxels along the op's x, y, z */
    float center[3];   /* centre of the image box, mm */
    float size[3];     /* full extent of the box, mm (voxel pitch = size/shape) */
    int64_t stride[3]; /* elements per step along x, y, z; {0,0,0} = contiguous [gz][gy][gx] */
} srfx_grid;

typedef struct srfx_tor_model {
    float kernel_width;      /* attr kernel_width: sigma_tor = kernel_width / 6 */
    int32_t geo_sigma2_flag; /* attr geo_sigma2_flag: scale sigma_tor^2 by the per-LOR factor */
    float tof_bin;           /* attr tof_bin, mm */
    float tof_sigma2;        /* attr tof_sigma2, mm^2; >= 405758 switches TOF off */
} srfx_tor_model;

/* ---------------------------------------------------------------- TOR --- */

/* Replaces op "Projection" of tor.so: tor.cc:14-28 (registration), 106-181
 * (Compute), tor.cu.cc:348-392 (host loop over slices), 244-278 (kernel).
 * lors: 10 columns x1,y1,z1,x2,y2,z2,xc,yc,zc,sigma2_factor in the op frame.
 * proj[e] = sum over strictly crossed slices and the PxP patch of
 *           w_TOR * w_TOF * image[voxel]. */
int srfx_tor_project(const float *lors, int64_t n, int64_t ld, const float *image,
                     srfx_grid grid, srfx_tor_model model, float *proj, uint32_t flags,
                     void *stream);

/* Replaces op "Backprojection" of tor.so: tor.cc:30-47, 203-271;
 * tor.cu.cc:394-433, 283-315, 169-209.
 * image[voxel] += w / lors_value[e] for every event with lors_value[e] > 1e-7
 * (the list-mode MLEM ratio folded into the scatter, A^T(1/v)). */
int srfx_tor_backproject(const float *lors, int64_t n, int64_t ld, const float *lors_value,
                         srfx_grid grid, srfx_tor_model model, float *image, uint32_t flags,
                         void *stream);

/* Plain adjoint image[voxel] += w * g[e], signed g, no threshold.  No
 * counterpart in the reference (its backprojection divides by v); it is the
 * backward pass of the differentiable projector. */
int srfx_tor_backproject_adjoint(const float *lors, int64_t n, int64_t ld, const float *g,
                                 srfx_grid grid, srfx_tor_model model, float *image,
                                 uint32_t flags, void *stream);

/* Replaces op "Maplors" of tor.so: tor.cc:49-64, 290-348; tor.cu.cc:435-481,
 * 316-346, 212-242, 94-122.  lors: 7 columns x1..z2, sigma2_factor.
 * image[voxel] += w_TOR / lors_value[e] (no TOF, no threshold). */
int srfx_tor_maplors(const float *lors, int64_t n, int64_t ld, const float *lors_value,
                     srfx_grid grid, float kernel_width, int32_t geo_sigma2_flag, float *image,
                     uint32_t flags, void *stream);

/* Fused list-mode OSEM subset pass for one axis group: for every event, the
 * projection p_e is formed in registers and, when p_e > 1e-7, w / p_e is
 * scattered into bp.  Equals srfx_tor_project followed by
 * srfx_tor_backproject(lors_value = proj) (model/recon_step.py:33-41), without
 * writing or re-reading p or re-reading the LORs. */
int srfx_tor_osem_subset(const float *lors, int64_t n, int64_t ld, const float *image,
                         srfx_grid image_grid, srfx_grid bp_grid, srfx_tor_model model,
                         float *bp, uint32_t flags, void *stream);

/* Number of non-zero (event, voxel) pairs of the TOR model: strictly crossed
 * slice, voxel inside the grid and the patch, d^2 < 9 sigma_tor^2.
 * nnz_per_event (nullable, n int32) and nnz_total (nullable, one device
 * uint64, added to) receive the counts. */
int srfx_tor_count_nnz(const float *lors, int64_t n, int64_t ld, srfx_grid grid,
                       srfx_tor_model model, int32_t *nnz_per_event, uint64_t *nnz_total,
                       void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): srfx_tor_osem_subset under a slab schedule (the flags'
 * SRFX_SLAB_* field, SRFX_SLAB_XCD8 when it names none), with the per-slab
 * partial projections in a static scratch of the library (16 MiB per device,
 * 2^22 floats, kept until the process ends) instead of a workspace from the
 * device's memory pool.  Calls no allocating routine.  Limit: slabs x n <=
 * 2^22 (n <= 524288 events at 8 slabs, 262144 at 16); larger inputs return
 * SRFX_ETOOLARGE before any device work (split them and pass
 * SRFX_ACCUMULATE).  Every partial is written before it is read, so nothing
 * needs clearing between calls; calls that overlap on one device (other
 * streams or host threads) share the scratch and race. */
int srfx_tor_osem_subset_noalloc(const float *lors, int64_t n, int64_t ld, const float *image,
                                 srfx_grid image_grid, srfx_grid bp_grid, srfx_tor_model model,
                                 float *bp, uint32_t flags, void *stream);

/* The reference's TOR patch width for a grid and kernel width, P =
 * ceil((sqrt(2) kernel_width + inter_z) / inter_x) (tor.cu.cc:369-375), and
 * the candidate box the kernels compile for it: 3 or 4 (NB x NB voxels per
 * slice hold every kept voxel of an event whose slice axis is its dominant
 * direction), or 0 when the ellipse of such an event can span more than 4
 * voxels (the generic chord walk).  Host only: no device work.  Either
 * output pointer may be NULL.  SRFX_EUNSUPPORTED for P > 16, as the ops. */
int srfx_tor_patch_info(srfx_grid grid, float kernel_width, int32_t *patch, int32_t *box);

/* ------------------------------------------------------------- Siddon --- */

/* Replaces op "Projection" of siddon.so: siddon.cc:14-26, 88-136;
 * siddon.cu.cc:583-606, 424-471, 293-331.  lors: 7 columns x1..z2, tof.
 * proj[e] = sum over traversed voxels of image * l / L^2 [* TOF weight]. */
int srfx_siddon_project(const float *lors, int64_t n, int64_t ld, const float *image,
                        srfx_grid grid, float tof_bin, float tof_sigma2, float *proj,
                        uint32_t flags, void *stream);

/* Replaces op "Backprojection" of siddon.so: siddon.cc:28-43, 152-205;
 * siddon.cu.cc:614-641, 480-526, 339-381.  image += a / v for v > 0. */
int srfx_siddon_backproject(const float *lors, int64_t n, int64_t ld, const float *lors_value,
                            srfx_grid grid, float tof_bin, float tof_sigma2, float *image,
                            uint32_t flags, void *stream);

/* Plain adjoint image += a * g (no reference counterpart; autograd backward). */
int srfx_siddon_backproject_adjoint(const float *lors, int64_t n, int64_t ld, const float *g,
                                    srfx_grid grid, float tof_bin, float tof_sigma2,
                                    float *image, uint32_t flags, void *stream);

/* Replaces op "Maplors" of siddon.so: siddon.cc:45-57, 221-268;
 * siddon.cu.cc:650-675, 535-575, 389-415.  Non-TOF, weight 1/(L*L),
 * image += a / v for v > 0.  lors: at least 6 columns. */
int srfx_siddon_maplors(const float *lors, int64_t n, int64_t ld, const float *lors_value,
                        srfx_grid grid, float *image, uint32_t flags, void *stream);

/* Fused Siddon OSEM subset pass: project, then scatter a / p for p > 0. */
int srfx_siddon_osem_subset(const float *lors, int64_t n, int64_t ld, const float *image,
                            srfx_grid grid, float tof_bin, float tof_sigma2, float *bp,
                            uint32_t flags, void *stream);

/* Non-zero Siddon segments (l > 0, and t^2/sigma^2 < 9 when TOF is on). */
int srfx_siddon_count_nnz(const float *lors, int64_t n, int64_t ld, srfx_grid grid,
                          float tof_bin, float tof_sigma2, int32_t *nnz_per_event,
                          uint64_t *nnz_total, void *stream);

/* ------------------------------------------------------------- Tiled --- */

/* Tiled TOR ops over a fixed event set (no counterpart in the reference,
 * whose ops re-test every event against every slice, tor.cu.cc:377-391; they
 * compute what srfx_tor_project / _backproject / _backproject_adjoint /
 * _maplors / _osem_subset compute for the same events).  A plan bins each
 * event into the tiles of `slices` x `wx` x `wy` op-frame voxels it can
 * touch; scatters then accumulate per tile in LDS and flush each tile with
 * coalesced atomics, and the tiled projection gathers from tiles staged in
 * LDS.  `chunk` caps the events one workgroup takes from a tile; `block` is
 * the workgroup size (64, 128, 256; 0 = 256).  slices * wx * wy <= 8192
 * (the scatter's tile is FP64 in LDS, 64 KB at most).
 * The tiled projection (srfx_tor_plan_project, SRFX_TILED_PROJECTION) needs a
 * plan built with `projection` set (8 more bytes per listing). */
typedef struct srfx_tile {
    int32_t slices, wx, wy, chunk, block;
    int32_t projection; /* nonzero: keep the tables of the tiled projection */
} srfx_tile;

typedef struct srfx_tor_plan_s *srfx_tor_plan;

/* Build a plan for n events (ncols = 10: TOF rows as on_tor_lors.py
 * produces them; ncols = 7: x1..z2 + sigma2 factor, for Maplors) of one
 * group frame.  Keeps the lors pointer: the events must stay unchanged and
 * alive while the plan is used.  Allocates the plan's tables and synchronises
 * `stream`. */
int srfx_tor_plan_create(const float *lors, int64_t n, int64_t ld, int32_t ncols,
                         srfx_grid grid, srfx_tor_model model, srfx_tile tile,
                         srfx_tor_plan *plan, void *stream);

/* The plans of consecutive event ranges of one LOR array, built together:
 * plans[k] is the plan srfx_tor_plan_create would build for events
 * bounds[k] .. bounds[k + 1] - 1 (bounds: nplans + 1 non-decreasing host
 * values, 0 <= bounds[0], bounds[nplans] <= n), e.g. the OSEM subsets of one
 * axis group (contiguous rows, graph/reconstruction/worker.py:79-85).  One
 * count pass and one fill pass over all events, two host synchronisations in
 * all and two shared device blocks, instead of per-plan passes and
 * allocations; each plan is destroyed on its own.  No reference
 * counterpart. */
int srfx_tor_plan_create_subsets(const float *lors, int64_t n, int64_t ld, int32_t ncols,
                                 srfx_grid grid, srfx_tor_model model, srfx_tile tile,
                                 int32_t nplans, const int64_t *bounds, srfx_tor_plan *plans,
                                 void *stream);

/* One LOR array of srfx_tor_plan_create_subset_groups: the arguments
 * srfx_tor_plan_create_subsets takes for it. */
typedef struct srfx_subset_group {
    const float *lors;
    int64_t n, ld;
    srfx_grid grid;
    srfx_tile tile; /* projection must be 0 */
    int32_t nplans;
    const int64_t *bounds; /* nplans + 1 host values */
} srfx_subset_group;

/* srfx_tor_plan_create_subsets for several LOR arrays (e.g. the x, y and z
 * groups of one event set) in one build: plans receives group 0's nplans
 * plans, then group 1's, and so on, each as srfx_tor_plan_create_subsets
 * would build it.  All groups share the count and fill phases, their two
 * host synchronisations and one chunk sort launch.  Groups with
 * tile.projection set are refused (SRFX_EINVAL).  No reference
 * counterpart. */
int srfx_tor_plan_create_subset_groups(const srfx_subset_group *groups, int32_t ngroups,
                                       int32_t ncols, srfx_tor_model model,
                                       srfx_tor_plan *plans, void *stream);

/* Listings (event, tile pairs) and workgroup chunks of a plan. */
int srfx_tor_plan_info(srfx_tor_plan plan, int64_t *nlistings, int32_t *nchunks);

/* The candidate box of the plan's walks (3 or 4; 0: chord walks) and
 * whether every event walks the NB x (NB - 1) box.  With box 3, such a plan
 * scatters through the kernel that holds only that walk when its tile takes
 * at most a fifth of a CU's LDS.  No reference counterpart. */
int srfx_tor_plan_walks(srfx_tor_plan plan, int32_t *box, int32_t *single_walk);

/* Synchronises the device, then frees the plan. */
int srfx_tor_plan_destroy(srfx_tor_plan plan);

/* Projection (tor.cc:16-28) of the plan's events; deterministic.  Plans
 * built with `projection` use the tiled gather; the others (and every call
 * with SRFX_IMAGE_OP_FRAME, whose image is the op-frame copy) walk each
 * event from its walk record, as the first kernel of the plan's OSEM step
 * does, with the flags' schedule. */
int srfx_tor_plan_project(srfx_tor_plan plan, const float *image, float *proj, uint32_t flags,
                          void *stream);

/* Backprojection (tor.cc:30-47), adjoint, Maplors (tor.cc:49-64; ncols 7
 * plans) into an image of the plan's grid. */
int srfx_tor_plan_backproject(srfx_tor_plan plan, const float *lors_value, float *image,
                              uint32_t flags, void *stream);
int srfx_tor_plan_backproject_adjoint(srfx_tor_plan plan, const float *g, float *image,
                                      uint32_t flags, void *stream);
int srfx_tor_plan_maplors(srfx_tor_plan plan, const float *lors_value, float *image,
                          uint32_t flags, void *stream);

/* OSEM subset step of the plan's events, as srfx_tor_osem_subset: the
 * projection (untiled kernel with the flags' schedule, or tiled with
 * SRFX_TILED_PROJECTION) into a plan buffer, then the tiled backprojection
 * of 1/p into bp.  Calls on one plan must not overlap. */
int srfx_tor_plan_osem_subset(srfx_tor_plan plan, const float *image, srfx_grid bp_grid,
                              float *bp, uint32_t flags, void *stream);

/* The backprojection scales the last OSEM step's projection stored in the
 * plan's event records, out[e] = 1 / p_e for p_e > 1e-7, else 0 (the
 * reference's ratio threshold, tor.cu.cc:203-206), in event order (n floats).
 * Lets a caller read the projection of the untiled OSEM step without a
 * second projection.  No reference counterpart (the reference's projection
 * output is the tensor of op "Projection", tor.cc:122-128). */
int srfx_tor_plan_scales(srfx_tor_plan plan, float *out, void *stream);

/* Tiled Siddon backprojection over a fixed LOR set (no counterpart in the
 * reference, whose BackRayTracing adds every segment to the image with its
 * own atomic, siddon.cu.cc:339-381; the ops compute what
 * srfx_siddon_backproject / _backproject_adjoint / _osem_subset compute for
 * the same LORs, with the same voxels and segment values, summed in another
 * order).  A plan cuts each LOR's walk into runs, one per image tile of
 * tx x ty x tz voxels it crosses (tx * ty * tz <= 8192: the scatter's tile is
 * FP64 in LDS), and lists the runs by tile; a scatter workgroup takes up to
 * `chunk` runs of one tile (0 = 2048), accumulates them in LDS and flushes
 * the tile with one atomic per non-zero voxel.  `block`: 64, 128 or 256
 * (0 = 256).  `sort` nonzero: the plan keeps a copy of its LORs sorted by
 * direction and position (a radix sort on the device), which its passes and
 * its OSEM projection read, so that the lanes of a wavefront trace nearly
 * parallel rays through neighbouring voxels; values and outputs stay in the
 * caller's LOR order.  Grid dimensions below 65536. */
typedef struct srfx_siddon_tile {
    int32_t tx, ty, tz, chunk, block, sort;
} srfx_siddon_tile;

typedef struct srfx_siddon_plan_s *srfx_siddon_plan;

/* Build a plan for n LORs (7 columns x1..z2, tof, as srfx_siddon_project) of
 * an image grid (the op's read and written images share its strides).  Keeps
 * the lors pointer: the LORs must stay unchanged and alive while the plan is
 * used.  Allocates the plan's tables and synchronises `stream`. */
int srfx_siddon_plan_create(const float *lors, int64_t n, int64_t ld, srfx_grid grid,
                            float tof_bin, float tof_sigma2, srfx_siddon_tile tile,
                            srfx_siddon_plan *plan, void *stream);

/* The plans of consecutive LOR ranges of one LOR array, built together:
 * plans[k] is the plan srfx_siddon_plan_create would build for LORs
 * bounds[k] .. bounds[k + 1] - 1 (bounds: nplans + 1 non-decreasing host
 * values, 0 <= bounds[0], bounds[nplans] <= n), e.g. the OSEM subsets
 * (contiguous rows, graph/reconstruction/worker.py:79-85).  One count pass
 * and one fill pass over all LORs, two host synchronisations and two shared
 * device blocks; each plan is destroyed on its own. */
int srfx_siddon_plan_create_subsets(const float *lors, int64_t n, int64_t ld, srfx_grid grid,
                                    float tof_bin, float tof_sigma2, srfx_siddon_tile tile,
                                    int32_t nplans, const int64_t *bounds,
                                    srfx_siddon_plan *plans, void *stream);

/* Runs (LOR, tile pairs) and workgroup chunks of a plan. */
int srfx_siddon_plan_info(srfx_siddon_plan plan, int64_t *nruns, int32_t *nchunks);

/* Synchronises the device, then frees the plan. */
int srfx_siddon_plan_destroy(srfx_siddon_plan plan);

/* Backprojection image += a / v for v > 0 (siddon.cc:28-43) and the plain
 * adjoint image += a * g of the plan's LORs. */
int srfx_siddon_plan_backproject(srfx_siddon_plan plan, const float *lors_value, float *image,
                                 uint32_t flags, void *stream);
int srfx_siddon_plan_backproject_adjoint(srfx_siddon_plan plan, const float *g, float *image,
                                         uint32_t flags, void *stream);

/* OSEM subset step of the plan's LORs, as srfx_siddon_osem_subset: the
 * projection (srfx_siddon_project, over the sorted LOR copy when the plan
 * has one) into a plan buffer, then the tiled backprojection of 1/p for
 * p > 0 into bp.  With counts (n floats, the caller's LOR order; NULL = all
 * ones) the scale is c/p, the backprojection of lors_value = p / c that
 * list-mode data with counts feeds the reference's backprojection
 * (siddon.cu.cc:363-371).  SRFX_OSEM_PROJECT_ONLY / SRFX_OSEM_SCATTER_ONLY
 * run one half (the projection stays in the plan in between).  Calls on one
 * plan must not overlap. */
int srfx_siddon_plan_osem_subset(srfx_siddon_plan plan, const float *image,
                                 const float *counts, float *bp, uint32_t flags, void *stream);

/* The projection of the last srfx_siddon_plan_osem_subset call (n floats, LOR
 * order), copied to out.  No reference counterpart. */
int srfx_siddon_plan_projection(srfx_siddon_plan plan, float *out, void *stream);

/* ------------------------------------------------------------ Batched --- */

/* One list-mode TOF reconstruction of a batch.  The image, the inverse
 * efficiency map and the backprojection buffer are C-order (nx, ny, nz)
 * float32 arrays (the reference's L3 image layout, data/image.py:4-28).  The
 * three LOR groups are 10-row SoA arrays in their group frames, as
 * on_tor_lors.py:251-262 produces them (op frame of group x = world (y, z, x),
 * group y = (x, z, y), group z = (x, y, z); physics/split_lors.py:54-85). */
typedef struct srfx_recon {
    int32_t shape[3];
    float center[3];
    float size[3];
    float *image;
    const float *emap;
    float *bp;
    const float *lors[3];
    int64_t n[3];
    int64_t ld[3];
} srfx_recon;

typedef struct srfx_tor_batch srfx_tor_batch;

/* Plan a batch of independent OSEM reconstructions that share one TOR model
 * and one subset count.  Allocates the plan's device tables; no reference
 * counterpart (the reference reconstructs one image per process,
 * graph/reconstruction/local_reconstruction.py:44-51).  SRFX_EUNSUPPORTED for
 * an image of 2^31 bytes or more, or with an op-frame slice of 2^24 voxels or
 * more (the frame gathers' 32-bit offsets). */
int srfx_tor_batch_create(const srfx_recon *recons, int32_t nrecon, int32_t nb_subsets,
                          srfx_tor_model model, srfx_tor_batch **plan);

/* Advance every reconstruction of the plan by OSEM subset `subset`
 * (ReconStep + mlem_update, model/recon_step.py:33-49, on contiguous subset
 * rows, graph/reconstruction/worker.py:79-85) in one persistent launch.
 * Enqueues kernels only (no allocation, no memset): capturable in a HIP
 * graph.  Calls on one plan must not overlap. */
int srfx_tor_batch_subset(srfx_tor_batch *plan, int32_t subset, void *stream);

/* Schedule of the plan's persistent projection: lanes per event (1, 4 or 16)
 * and events per work item a wave pulls (64, 128 or 256); default 4 and 64.
 * SRFX_EINVAL for any other pair (the schedule is then unchanged).  Set it
 * before capturing subsets in a HIP graph: a graph keeps the kernel it
 * recorded.  No reference counterpart (launch tuning of the batch). */
int srfx_tor_batch_set_schedule(srfx_tor_batch *plan, int32_t lanes_per_event,
                                int32_t events_per_item);

int srfx_tor_batch_destroy(srfx_tor_batch *plan);

/* ------------------------------------------------------------- Update --- */

/* mlem_update, model/recon_step.py:48-49: x <- x * emap * bp (elementwise,
 * n voxels, evaluated as (x * emap) * bp). */
int srfx_mlem_update(float *x, const float *emap, const float *bp, int64_t n, void *stream);

/* Copy of an image into an op frame, the reference's transpose of the image
 * into an axis group's frame (split_lors.py:79): out[k][j][i] = img[i * sx +
 * j * sy + k * sz] with (sx, sy, sz) the element strides of grid (all zero:
 * already contiguous), out contiguous [gz][gy][gx].  img and out must not
 * overlap.  The projections then read voxel rows from shared cache lines
 * (SRFX_IMAGE_OP_FRAME, or a contiguous grid). */
int srfx_frame_copy(const float *img, srfx_grid grid, float *out, void *stream);

/* PSF blur around the projector (PSFReconStep, model/recon_step.py:56-115).
 * The image, nx x ny x nz contiguous with z fastest, is a matrix X of nrows =
 * nx * ny rows by nz columns; out = S X M, S a sparse nrows x nrows matrix in
 * CSR (row_ptr [nrows + 1], col / val [row_ptr[nrows]], 0 <= col < nrows),
 * M dense nz x nz row-major, zero outside the band z - below <= z' <= z +
 * above of each column z (below = above = nz - 1 for any M; the entries
 * outside the band are not read).  The reference's forward blur
 * (recon_step.py:79-90) is S = psf_xy, M = psf_z^T; its backward blur
 * (103-110) is S = psf_xy^T, M = psf_z.  img and out must not overlap;
 * nz <= 2048. */
int srfx_psf_apply(const int32_t *row_ptr, const int32_t *col, const float *val, int32_t nrows,
                   const float *m, int32_t nz, int32_t below, int32_t above, const float *img,
                   float *out, void *stream);

/* Status text for a return code. */
const char *srfx_strerror(int status);

#ifdef __cplusplus
}
#endif

#endif /* SRFX_H */
