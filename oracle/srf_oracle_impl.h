/*
 * TEST INFRASTRUCTURE ONLY — not part of the product.  Only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load the
 * library built from this file; the srf_amd package never does.
 *
 * CPU restatement of the reference projector kernels.  This file is included
 * twice by srf_oracle.c:
 *   REAL = double  -> the FP64 oracle (every intermediate in double);
 *   REAL = float   -> the FP32 "port": float variables where the reference has
 *                     float variables, double arithmetic where the reference's
 *                     double literals promote (std::exp(-0.5 * ...), 2.0*M_PI,
 *                     0.5*(...), 9.0*...).  This is the CPU baseline.
 *
 * Sources restated (paths relative to the reference repository root):
 *   src/cpp/tf_custom_ops/tor.cu.cc     TOR projection / backprojection / map
 *   src/cpp/tf_custom_ops/siddon.cu.cc  Siddon projection / backprojection / map
 * Each function below cites the lines it follows.
 *
 * One deliberate divergence, shared with the HIP kernels: a Siddon ray whose
 * start voxel lies outside the grid (possible only for a ray lying exactly on
 * the box's upper face along an axis it does not move along, or for non-finite
 * input) is skipped; the reference indexes out of bounds there
 * (siddon.cu.cc:235-243 leaves pos == grid).  The traversal is also capped at
 * gx+gy+gz+3 steps, which a finite ray never reaches.
 */

#ifndef REAL
#error "define REAL before including srf_oracle_impl.h"
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUFFIX)

/* ------------------------------------------------------------------------ */
/* TOR (tor.cu.cc)                                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    int gx, gy, gz;
    REAL inter_x, inter_y, inter_z;
    REAL l_bound, b_bound;
    REAL tor_sigma2;
    int patch;
    REAL cz, lz;
} FN(TorConst);

/* Host-side constants: tor.cu.cc:364-375 (projection), 411-418 (bp),
 * 449-462 (map).  kernel_width*kernel_width/36 is float arithmetic in the
 * reference; the patch size is evaluated in double (std::sqrt(2) promotes). */
static void FN(tor_setup)(const int *grid, const float *center, const float *size,
                          float kernel_width, FN(TorConst) * k)
{
    k->gx = grid[0];
    k->gy = grid[1];
    k->gz = grid[2];
    REAL lx = size[0], ly = size[1], lz = size[2];
    k->inter_x = lx / (REAL)(unsigned)grid[0];
    k->inter_y = ly / (REAL)(unsigned)grid[1];
    k->inter_z = lz / (REAL)(unsigned)grid[2];
    k->l_bound = (REAL)center[0] - lx / 2;
    k->b_bound = (REAL)center[1] - ly / 2;
    REAL kw = kernel_width;
    k->tor_sigma2 = kw * kw / 36;
    k->patch = (int)ceil((sqrt(2.0) * (double)kw + (double)k->inter_z) / (double)k->inter_x);
    k->cz = center[2];
    k->lz = lz;
}

/* Slice centre along the slice axis.  Projection (tor.cu.cc:380) evaluates it
 * in float; backprojection and map (423, 471) divide by the double literal 2.0,
 * so the sum is formed in double and rounded once. */
static REAL FN(tor_slice_z)(const FN(TorConst) * k, unsigned s, int double_form)
{
    REAL step = (REAL)s * k->inter_z;
    if (double_form)
        return (REAL)((double)k->cz - (double)(k->lz - k->inter_z) / 2.0 + (double)step);
    return k->cz - (k->lz - k->inter_z) / 2 + step;
}

/* CalculateCrossPoint, tor.cu.cc:22-48.  Strict 0 < ratio < 1. */
static int FN(tor_cross)(REAL x1, REAL y1, REAL z1, REAL x2, REAL y2, REAL z2,
                         REAL slice_z, REAL *dcos_x, REAL *dcos_y, REAL *cross_x, REAL *cross_y)
{
    REAL d0 = x2 - x1, d1 = y2 - y1, d2 = z2 - z1;
    REAL ratio = (slice_z - z1) / d2;
    int flag = ratio > 0.0 && ratio < 1.0;
    *cross_x = x1 + d0 * ratio;
    *cross_y = y1 + d1 * ratio;
    REAL dis = (REAL)sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    *dcos_x = d0 / dis;
    *dcos_y = d1 / dis;
    return flag;
}

/* CalculateSMV (tor.cu.cc:50-92) when use_tof, MapSMV (94-122) otherwise.
 * Returns the weight; *inside reports d2 < 9 sigma^2 (the non-zero test). */
static REAL FN(tor_smv)(REAL xc, REAL yc, REAL zc, REAL f, int geo, REAL slice_z, REAL tof_bin,
                        REAL tof_sigma2, int use_tof, REAL cross_x, REAL cross_y, REAL mesh_x,
                        REAL mesh_y, REAL dcos_x, REAL dcos_y, REAL tor_sigma2, int *inside)
{
    REAL delta_x = cross_x - mesh_x;
    REAL delta_y = cross_y - mesh_y;
    REAL r_cos = (delta_x * dcos_x + delta_y * dcos_y);
    REAL d2 = delta_x * delta_x + delta_y * delta_y - r_cos * r_cos;
    REAL tsc = tor_sigma2;
    if (geo)
        tsc = tor_sigma2 * f;
    *inside = d2 < 9.0 * tsc;
    REAL value = *inside ? (REAL)exp(-0.5 * d2 / tsc) : (REAL)0.0;
    if (use_tof) {
        REAL d2_tof = ((xc - cross_x) * (xc - cross_x) + (yc - cross_y) * (yc - cross_y) +
                       (zc - slice_z) * (zc - slice_z) - d2);
        REAL tse = tof_sigma2 + (tof_bin * tof_bin) / 12;
        REAL t2 = d2_tof / tse;
        value *= tof_bin * exp(-0.5 * t2) / sqrt(2.0 * M_PI * tof_sigma2);
    }
    return value;
}

/* One event, all slices: the body of ComputeSlice / BackComputeSlice /
 * Mapping (tor.cu.cc:244-346) with LoopPatch / BackLoopPatch / MapLoopPatch
 * (124-242) inlined.  The reference launches one kernel per slice in
 * ascending order; per event this is the same slice order.
 *   mode 0: projection, returns sum w*x (accumulated in REAL).
 *   mode 1: backprojection, img[v] += w / v_i when v_i > 1e-7.
 *   mode 2: map (no TOF, float mesh formula), img[v] += w / v_i, no threshold.
 *   mode 3: plain adjoint, img[v] += w * g_i (no reference counterpart).
 *   mode 4: count non-zero (event, voxel) pairs. */
static REAL FN(tor_event)(const FN(TorConst) * k, const float *lors, int64_t ld, int64_t e,
                          int ncols, int geo, REAL tof_bin, REAL tof_sigma2, int mode,
                          const float *img, REAL *acc, REAL val, int64_t *nnz)
{
    REAL x1 = lors[0 * ld + e], y1 = lors[1 * ld + e], z1 = lors[2 * ld + e];
    REAL x2 = lors[3 * ld + e], y2 = lors[4 * ld + e], z2 = lors[5 * ld + e];
    REAL xc = 0, yc = 0, zc = 0, f;
    if (ncols == 10) {
        xc = lors[6 * ld + e];
        yc = lors[7 * ld + e];
        zc = lors[8 * ld + e];
        f = lors[9 * ld + e];
    } else {
        f = lors[6 * ld + e];
    }
    int use_tof = (mode != 2) && (tof_sigma2 < (REAL)405758.0f);
    int double_slice = (mode != 0 && mode != 4);
    unsigned slice_mesh_num = (unsigned)k->gx * (unsigned)k->gy;
    REAL pv = 0;
    for (unsigned s = 0; s < (unsigned)k->gz; s++) {
        REAL slice_z = FN(tor_slice_z)(k, s, double_slice);
        REAL dcos_x, dcos_y, cross_x, cross_y;
        if (!FN(tor_cross)(x1, y1, z1, x2, y2, z2, slice_z, &dcos_x, &dcos_y, &cross_x, &cross_y))
            continue;
        unsigned offset = s * slice_mesh_num;
        int ix0 = (int)((cross_x - k->l_bound) / k->inter_x) - (int)((unsigned)k->patch / 2);
        int iy0 = (int)((cross_y - k->b_bound) / k->inter_y) - (int)((unsigned)k->patch / 2);
        for (int j = 0; j < k->patch; j++) {
            for (int i = 0; i < k->patch; i++) {
                int index0 = ix0 + i, index1 = iy0 + j;
                if (index0 < 0 || index0 >= k->gx || index1 < 0 || index1 >= k->gy)
                    continue;
                int index = index0 + index1 * k->gx;
                REAL mesh_x, mesh_y;
                if (mode == 2) { /* MapLoopPatch: float arithmetic, tor.cu.cc:235-236 */
                    mesh_x = k->inter_x * (REAL)(0.5 + index0) + k->l_bound;
                    mesh_y = k->inter_y * (REAL)(0.5 + index1) + k->b_bound;
                } else { /* LoopPatch: index + 0.5 is double, tor.cu.cc:159-160 */
                    mesh_x = (REAL)((double)k->inter_x * (index0 + 0.5) + (double)k->l_bound);
                    mesh_y = (REAL)((double)k->inter_y * (index1 + 0.5) + (double)k->b_bound);
                }
                int inside;
                REAL w = FN(tor_smv)(xc, yc, zc, f, geo, slice_z, tof_bin, tof_sigma2, use_tof,
                                     cross_x, cross_y, mesh_x, mesh_y, dcos_x, dcos_y,
                                     k->tor_sigma2, &inside);
                switch (mode) {
                case 0:
                    pv += w * (REAL)img[offset + index];
                    break;
                case 1:
                    if (val > 1e-7)
                        acc[offset + index] += w / val;
                    break;
                case 2:
                    acc[offset + index] += w / val;
                    break;
                case 3:
                    acc[offset + index] += w * val;
                    break;
                default:
                    *nnz += inside;
                    break;
                }
            }
        }
    }
    return pv;
}

/* ------------------------------------------------------------------------ */
/* Siddon (siddon.cu.cc)                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    REAL start[3], dir[3];
    REAL length, min_t, max_t;
} FN(Ray);

typedef struct {
    int boffs;
    REAL nextT[3], deltaT[3];
    int deltaBuf[3], inBuf[3];
} FN(RayCast);

/* CreateRay, siddon.cu.cc:61-81 */
static void FN(sid_create_ray)(REAL x1, REAL y1, REAL z1, REAL x2, REAL y2, REAL z2, FN(Ray) * r)
{
    REAL dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
    REAL length = (REAL)sqrt(dx * dx + dy * dy + dz * dz);
    r->length = length;
    r->start[0] = x1;
    r->start[1] = y1;
    r->start[2] = z1;
    r->dir[0] = dx / length;
    r->dir[1] = dy / length;
    r->dir[2] = dz / length;
    r->min_t = 0;
    r->max_t = length;
}

/* ClipRay, siddon.cu.cc:118-153 */
static int FN(sid_clip)(REAL toplim, REAL botlim, int comp, FN(Ray) * r)
{
    REAL p0 = r->start[comp], dp = r->dir[comp];
    REAL tmax = r->max_t, tmin = r->min_t, t;
    if (dp > 0.0) {
        t = (toplim - p0) / dp;
        if (t < tmax)
            tmax = t;
        t = (botlim - p0) / dp;
        if (t > tmin)
            tmin = t;
    } else if (dp < 0.0) {
        t = (botlim - p0) / dp;
        if (t < tmax)
            tmax = t;
        t = (toplim - p0) / dp;
        if (t > tmin)
            tmin = t;
    } else {
        if ((p0 < botlim) || (p0 > toplim))
            return 0;
    }
    r->min_t = tmin;
    r->max_t = tmax;
    return tmin < tmax;
}

/* IsThroughImage, siddon.cu.cc:163-183 */
static int FN(sid_through)(const int *grid, const REAL *size, const REAL *center, FN(Ray) * r)
{
    (void)grid;
    r->min_t = 0;
    r->max_t = r->length;
    for (int i = 0; i < 3; i++)
        if (!FN(sid_clip)(center[i] + size[i] / 2, center[i] - size[i] / 2, i, r))
            return 0;
    return r->min_t < r->max_t;
}

/* SetupRayCastComponent, siddon.cu.cc:196-250 */
static void FN(sid_setup_comp)(const int *grid, const REAL *size, const FN(Ray) * r, int comp,
                               FN(RayCast) * rc, int *pos)
{
    REAL p0 = r->start[comp], dirc = r->dir[comp], mint = r->min_t;
    REAL pt = mint * dirc + p0;
    REAL v_res = size[comp] / grid[comp];
    int v_count = grid[comp];
    if (dirc > 0.0) {
        *pos = (int)(pt / v_res);
        rc->nextT[comp] = ((*pos + 1) * v_res - pt) / dirc;
        rc->deltaT[comp] = v_res / dirc;
        rc->deltaBuf[comp] = 1;
        rc->inBuf[comp] = v_count - *pos;
    } else if (dirc < 0.0) {
        *pos = v_count - 1 - (int)(v_count - pt / v_res);
        rc->nextT[comp] = (*pos * v_res - pt) / dirc;
        rc->deltaT[comp] = -v_res / dirc;
        rc->deltaBuf[comp] = -1;
        rc->inBuf[comp] = *pos + 1;
    } else {
        *pos = (int)(pt / v_res);
        rc->nextT[comp] = REAL_MAX;
        rc->deltaT[comp] = REAL_MAX;
        rc->deltaBuf[comp] = 0;
        rc->inBuf[comp] = 1;
    }
    for (int i = comp - 1; i >= 0; i--)
        rc->deltaBuf[comp] *= grid[i];
}

/* SetupRayCast, siddon.cu.cc:261-284.  Returns 0 when a start index is
 * outside the grid (the guarded divergence described in the header). */
static int FN(sid_setup)(const int *grid, const REAL *size, const REAL *center, FN(Ray) * r,
                         FN(RayCast) * rc)
{
    for (int i = 0; i < 3; i++)
        r->start[i] = r->start[i] - (center[i] - size[i] / 2);
    int pos[3] = {0, 0, 0};
    for (int c = 0; c < 3; c++)
        FN(sid_setup_comp)(grid, size, r, c, rc, &pos[c]);
    rc->boffs = pos[0] + grid[0] * (pos[1] + grid[1] * pos[2]);
    for (int c = 0; c < 3; c++)
        if (pos[c] < 0 || pos[c] >= grid[c])
            return 0;
    return 1;
}

/* One Siddon event: project (siddon.cu.cc:424-471 + RayTracing 293-331),
 * backproject (480-526 + BackRayTracing 339-381), map (535-575 + Map 389-415).
 *   mode 0: projection; mode 1: bp, += a/v when v > 0; mode 2: map (no TOF,
 *   weight 1/(L*L)), += a/v when v > 0; mode 3: plain adjoint += a*g;
 *   mode 4: count non-zero segments. */
static REAL FN(sid_event)(const int *grid, const REAL *size, const REAL *center, const float *lors,
                          int64_t ld, int64_t e, REAL tof_bin, REAL tof_sigma2, int mode,
                          const float *img, REAL *acc, REAL val, int64_t *nnz)
{
    FN(Ray) ray;
    FN(sid_create_ray)(lors[0 * ld + e], lors[1 * ld + e], lors[2 * ld + e], lors[3 * ld + e],
                       lors[4 * ld + e], lors[5 * ld + e], &ray);
    REAL weight;
    if (mode == 2)
        weight = 1 / (ray.length * ray.length); /* siddon.cu.cc:557 */
    else
        weight = 1 / ray.length / ray.length; /* siddon.cu.cc:459, 514 */
    if (!FN(sid_through)(grid, size, center, &ray))
        return 0;
    FN(RayCast) rc;
    if (!FN(sid_setup)(grid, size, center, &ray, &rc))
        return 0;
    REAL t_tof = 0;
    if (mode != 2)
        t_tof = (REAL)(ray.length * 0.5 - ray.min_t - (REAL)lors[6 * ld + e]);
    int use_tof = (mode != 2) && (tof_sigma2 < (REAL)405758.0f);
    REAL last_t = 0.0, l_len;
    int stillin = rc.inBuf[0] * rc.inBuf[1] * rc.inBuf[2];
    int cap = grid[0] + grid[1] + grid[2] + 3;
    REAL pv = 0;
    while (stillin && cap-- > 0) {
        int i = (rc.nextT[0] <= rc.nextT[1]) ? ((rc.nextT[0] <= rc.nextT[2]) ? 0 : 2)
                                             : ((rc.nextT[1] <= rc.nextT[2]) ? 1 : 2);
        l_len = rc.nextT[i] - last_t;
        int mi = rc.boffs;
        if (use_tof) {
            REAL bs = tof_bin;
            REAL t = (REAL)(t_tof - 0.5 * (rc.nextT[i] + last_t));
            REAL sigma2 = (REAL)(tof_sigma2 + (l_len * l_len + bs * bs) / 12.0);
            REAL t2 = t * t / sigma2;
            if (t2 < 9) {
                switch (mode) {
                case 0:
                    pv += (REAL)img[mi] * weight * l_len * tof_bin * exp(-0.5 * t2) /
                          sqrt(2.0 * M_PI * sigma2);
                    break;
                case 1:
                case 3: {
                    REAL value =
                        weight * l_len * tof_bin * exp(-0.5 * t2) / sqrt(2.0 * M_PI * sigma2);
                    if (mode == 1) {
                        if (val > 0)
                            acc[mi] += value / val;
                    } else {
                        acc[mi] += value * val;
                    }
                    break;
                }
                default:
                    *nnz += (l_len > 0);
                    break;
                }
            }
        } else {
            switch (mode) {
            case 0:
                pv += (REAL)img[mi] * weight * l_len;
                break;
            case 1:
            case 2:
                if (val > 0)
                    acc[mi] += weight * l_len / val;
                break;
            case 3:
                acc[mi] += weight * l_len * val;
                break;
            default:
                *nnz += (l_len > 0);
                break;
            }
        }
        last_t = rc.nextT[i];
        rc.nextT[i] += rc.deltaT[i];
        rc.boffs += rc.deltaBuf[i];
        stillin = rc.inBuf[i] - 1;
        rc.inBuf[i] = stillin;
    }
    return pv;
}

/* ------------------------------------------------------------------------ */
/* Exported drivers.  Outputs are zeroed first, as the reference ops zero      */
/* theirs with cudaMemset (tor.cc:132, 222, 310; siddon.cc memsets).  Scatter */
/* ops use one private image per thread and add them in thread order, so the  */
/* result is deterministic for a fixed thread count.                          */
/* ------------------------------------------------------------------------ */

static void FN(set_threads)(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* lors: 10 rows (x1,y1,z1,x2,y2,z2,xc,yc,zc,sigma2_factor) of ld floats
 * (tor.cc:140-160), or 7 rows (x1..z2, sigma2_factor) for the map
 * (tor.cc:314-320).  grid/center/size are in the group's frame. */
int FN(tor_project)(const float *lors, int64_t n, int64_t ld, const int *grid,
                    const float *center, const float *size, const float *img,
                    float kernel_width, int geo, float tof_bin, float tof_sigma2, int nthreads,
                    REAL *out)
{
    FN(TorConst) k;
    FN(tor_setup)(grid, center, size, kernel_width, &k);
    FN(set_threads)(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t e = 0; e < n; e++)
        out[e] = FN(tor_event)(&k, lors, ld, e, 10, geo, tof_bin, tof_sigma2, 0, img, NULL, 0,
                               NULL);
    return 0;
}

static int FN(tor_scatter)(const float *lors, int64_t n, int64_t ld, int ncols, const int *grid,
                           const float *center, const float *size, const float *vals,
                           float kernel_width, int geo, float tof_bin, float tof_sigma2,
                           int mode, int nthreads, REAL *out)
{
    FN(TorConst) k;
    FN(tor_setup)(grid, center, size, kernel_width, &k);
    int64_t nvox = (int64_t)grid[0] * grid[1] * grid[2];
    memset(out, 0, sizeof(REAL) * nvox);
    FN(set_threads)(nthreads);
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    REAL *priv = (REAL *)calloc((size_t)nt * nvox, sizeof(REAL));
    if (!priv)
        return 2;
#pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        REAL *mine = priv + (size_t)t * nvox;
#pragma omp for schedule(static)
        for (int64_t e = 0; e < n; e++)
            FN(tor_event)(&k, lors, ld, e, ncols, geo, tof_bin, tof_sigma2, mode, NULL, mine,
                          (REAL)vals[e], NULL);
    }
    for (int t = 0; t < nt; t++)
        for (int64_t v = 0; v < nvox; v++)
            out[v] += priv[(size_t)t * nvox + v];
    free(priv);
    return 0;
}

int FN(tor_backproject)(const float *lors, int64_t n, int64_t ld, const int *grid,
                        const float *center, const float *size, const float *lors_value,
                        float kernel_width, int geo, float tof_bin, float tof_sigma2,
                        int nthreads, REAL *out)
{
    return FN(tor_scatter)(lors, n, ld, 10, grid, center, size, lors_value, kernel_width, geo,
                           tof_bin, tof_sigma2, 1, nthreads, out);
}

int FN(tor_backproject_adjoint)(const float *lors, int64_t n, int64_t ld, const int *grid,
                                const float *center, const float *size, const float *g,
                                float kernel_width, int geo, float tof_bin, float tof_sigma2,
                                int nthreads, REAL *out)
{
    return FN(tor_scatter)(lors, n, ld, 10, grid, center, size, g, kernel_width, geo, tof_bin,
                           tof_sigma2, 3, nthreads, out);
}

int FN(tor_map)(const float *lors, int64_t n, int64_t ld, const int *grid, const float *center,
                const float *size, const float *lors_value, float kernel_width, int geo,
                int nthreads, REAL *out)
{
    return FN(tor_scatter)(lors, n, ld, 7, grid, center, size, lors_value, kernel_width, geo,
                           0.0f, 1.0e6f, 2, nthreads, out);
}

/* Per-event count of (event, voxel) pairs with d2 < 9 sigma^2 inside the grid
 * and the patch, over strictly crossed slices. */
int FN(tor_count)(const float *lors, int64_t n, int64_t ld, const int *grid,
                  const float *center, const float *size, float kernel_width, int geo,
                  int nthreads, int64_t *nnz_per_event)
{
    FN(TorConst) k;
    FN(tor_setup)(grid, center, size, kernel_width, &k);
    FN(set_threads)(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t e = 0; e < n; e++) {
        int64_t c = 0;
        FN(tor_event)(&k, lors, ld, e, 10, geo, 1.0f, 1.0e6f, 4, NULL, NULL, 0, &c);
        nnz_per_event[e] = c;
    }
    return 0;
}

/* lors: 7 rows (x1,y1,z1,x2,y2,z2,tof) of ld floats (siddon.cc:108-122). */
int FN(siddon_project)(const float *lors, int64_t n, int64_t ld, const int *grid,
                       const float *center, const float *size, const float *img, float tof_bin,
                       float tof_sigma2, int nthreads, REAL *out)
{
    REAL sz[3] = {size[0], size[1], size[2]}, ce[3] = {center[0], center[1], center[2]};
    FN(set_threads)(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t e = 0; e < n; e++)
        out[e] = FN(sid_event)(grid, sz, ce, lors, ld, e, tof_bin, tof_sigma2, 0, img, NULL, 0,
                               NULL);
    return 0;
}

static int FN(sid_scatter)(const float *lors, int64_t n, int64_t ld, const int *grid,
                           const float *center, const float *size, const float *vals,
                           float tof_bin, float tof_sigma2, int mode, int nthreads, REAL *out)
{
    REAL sz[3] = {size[0], size[1], size[2]}, ce[3] = {center[0], center[1], center[2]};
    int64_t nvox = (int64_t)grid[0] * grid[1] * grid[2];
    memset(out, 0, sizeof(REAL) * nvox);
    FN(set_threads)(nthreads);
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    REAL *priv = (REAL *)calloc((size_t)nt * nvox, sizeof(REAL));
    if (!priv)
        return 2;
#pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        REAL *mine = priv + (size_t)t * nvox;
#pragma omp for schedule(static)
        for (int64_t e = 0; e < n; e++)
            FN(sid_event)(grid, sz, ce, lors, ld, e, tof_bin, tof_sigma2, mode, NULL, mine,
                          (REAL)vals[e], NULL);
    }
    for (int t = 0; t < nt; t++)
        for (int64_t v = 0; v < nvox; v++)
            out[v] += priv[(size_t)t * nvox + v];
    free(priv);
    return 0;
}

int FN(siddon_backproject)(const float *lors, int64_t n, int64_t ld, const int *grid,
                           const float *center, const float *size, const float *lors_value,
                           float tof_bin, float tof_sigma2, int nthreads, REAL *out)
{
    return FN(sid_scatter)(lors, n, ld, grid, center, size, lors_value, tof_bin, tof_sigma2, 1,
                           nthreads, out);
}

int FN(siddon_backproject_adjoint)(const float *lors, int64_t n, int64_t ld, const int *grid,
                                   const float *center, const float *size, const float *g,
                                   float tof_bin, float tof_sigma2, int nthreads, REAL *out)
{
    return FN(sid_scatter)(lors, n, ld, grid, center, size, g, tof_bin, tof_sigma2, 3, nthreads,
                           out);
}

/* lors: at least 6 rows (siddon.cc:246-259); the tof row is not read. */
int FN(siddon_map)(const float *lors, int64_t n, int64_t ld, const int *grid, const float *center,
                   const float *size, const float *lors_value, int nthreads, REAL *out)
{
    return FN(sid_scatter)(lors, n, ld, grid, center, size, lors_value, 0.0f, 1.0e6f, 2,
                           nthreads, out);
}

int FN(siddon_count)(const float *lors, int64_t n, int64_t ld, const int *grid,
                     const float *center, const float *size, float tof_bin, float tof_sigma2,
                     int nthreads, int64_t *nnz_per_event)
{
    REAL sz[3] = {size[0], size[1], size[2]}, ce[3] = {center[0], center[1], center[2]};
    FN(set_threads)(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t e = 0; e < n; e++) {
        int64_t c = 0;
        FN(sid_event)(grid, sz, ce, lors, ld, e, tof_bin, tof_sigma2, 4, NULL, NULL, 0, &c);
        nnz_per_event[e] = c;
    }
    return 0;
}

#undef FN
#undef CAT
#undef CAT_
