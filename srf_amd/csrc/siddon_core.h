// Siddon ray setup and walk shared by the untiled kernels (siddon.hip) and
// the tiled backprojection (siddon_tile.hip).
//
// Semantics follow the reference CUDA op src/cpp/tf_custom_ops/siddon.cu.cc
// (CreateRay 61-81, ClipRay / IsThroughImage 118-183, SetupRayCast 196-284,
// RayTracing / BackRayTracing / Map 293-415).  Ray geometry and traversal
// (clip, start voxel, nextT updates, argmin, segment lengths and the TOF
// distance) are carried in FP64 with explicitly rounded operations in the
// reference's source order, so the visited voxels are those of the FP64
// restatement, including rays lying exactly on voxel faces, whose voxel is
// otherwise decided by FP32 rounding.  Weights (1/L^2, the TOF Gaussian) and
// the accumulation are FP32.
#pragma once

#include <float.h>
#include <math.h>

#include "common.h"

namespace srfx {
namespace {

enum SidMode { kProj = 0, kBack = 1, kMap = 2, kAdj = 3, kCount = 4, kOsem = 5 };

struct SidParams {
    const float *__restrict__ lors;
    int64_t n, ld;
    const float *__restrict__ val;
    const float *__restrict__ img;
    float *out;
    int32_t *nnz;
    unsigned long long *nnz_total;
    int g[3];
    float size[3], center[3];
    int s[3]; // read-image strides
    int o[3]; // written-image strides
    float tof_bin, tof_sigma2;
    int use_tof;
};

struct Walk {
    double nextT[3], deltaT[3];
    int inBuf[3], step[3];
    int pos[3];
    float weight;
    double t_tof;
    bool ok;
};

// Ray setup: CreateRay, IsThroughImage, SetupRayCast and the weight / TOF
// origin of project() (siddon.cu.cc:450-460).
template <int MODE>
__device__ __forceinline__ void setup_walk(const SidParams &p, int64_t e, Walk &w)
{
    const float *L = p.lors;
    const int64_t ld = p.ld;
    double P0[3], D[3];
    P0[0] = __builtin_nontemporal_load(L + 0 * ld + e);
    P0[1] = __builtin_nontemporal_load(L + 1 * ld + e);
    P0[2] = __builtin_nontemporal_load(L + 2 * ld + e);
    double x2 = __builtin_nontemporal_load(L + 3 * ld + e);
    double y2 = __builtin_nontemporal_load(L + 4 * ld + e);
    double z2 = __builtin_nontemporal_load(L + 5 * ld + e);
    double dx = x2 - P0[0], dy = y2 - P0[1], dz = z2 - P0[2]; // exact
    double length =
        __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz)));
    D[0] = __ddiv_rn(dx, length);
    D[1] = __ddiv_rn(dy, length);
    D[2] = __ddiv_rn(dz, length);
    if (MODE == kMap)
        w.weight = (float)(1.0 / (length * length));
    else
        w.weight = (float)(1.0 / length / length);
    w.ok = false;
    // IsThroughImage / ClipRay
    double tmin = 0.0, tmax = length;
    for (int c = 0; c < 3; c++) {
        double half = (double)p.size[c] / 2;
        double top = __dadd_rn((double)p.center[c], half);
        double bot = __dsub_rn((double)p.center[c], half);
        double p0 = P0[c], dp = D[c];
        if (dp > 0.0) {
            double t = __ddiv_rn(__dsub_rn(top, p0), dp);
            if (t < tmax)
                tmax = t;
            t = __ddiv_rn(__dsub_rn(bot, p0), dp);
            if (t > tmin)
                tmin = t;
        } else if (dp < 0.0) {
            double t = __ddiv_rn(__dsub_rn(bot, p0), dp);
            if (t < tmax)
                tmax = t;
            t = __ddiv_rn(__dsub_rn(top, p0), dp);
            if (t > tmin)
                tmin = t;
        } else {
            if ((p0 < bot) || (p0 > top))
                return;
        }
        if (!(tmin < tmax))
            return;
    }
    // SetupRayCast: origin at the box corner, per-axis start voxel.
    for (int c = 0; c < 3; c++) {
        double p0 =
            __dsub_rn(P0[c], __dsub_rn((double)p.center[c], (double)p.size[c] / 2));
        double dir = D[c];
        double pt = __dadd_rn(__dmul_rn(tmin, dir), p0);
        double v_res = __ddiv_rn((double)p.size[c], (double)p.g[c]);
        int vc = p.g[c];
        int pos;
        if (dir > 0.0) {
            pos = (int)__ddiv_rn(pt, v_res);
            w.nextT[c] = __ddiv_rn(__dsub_rn(__dmul_rn((double)(pos + 1), v_res), pt), dir);
            w.deltaT[c] = __ddiv_rn(v_res, dir);
            w.step[c] = 1;
            w.inBuf[c] = vc - pos;
        } else if (dir < 0.0) {
            pos = vc - 1 - (int)__dsub_rn((double)vc, __ddiv_rn(pt, v_res));
            w.nextT[c] = __ddiv_rn(__dsub_rn(__dmul_rn((double)pos, v_res), pt), dir);
            w.deltaT[c] = __ddiv_rn(-v_res, dir);
            w.step[c] = -1;
            w.inBuf[c] = pos + 1;
        } else {
            pos = (int)__ddiv_rn(pt, v_res);
            w.nextT[c] = DBL_MAX;
            w.deltaT[c] = DBL_MAX;
            w.step[c] = 0;
            w.inBuf[c] = 1;
        }
        if (pos < 0 || pos >= vc)
            return;
        w.pos[c] = pos;
    }
    if (MODE != kMap) {
        double tof = __builtin_nontemporal_load(L + 6 * ld + e);
        w.t_tof = __dsub_rn(__dsub_rn(__dmul_rn(length, 0.5), tmin), tof);
    } else {
        w.t_tof = 0.0;
    }
    w.ok = true;
}

// Walk the ray (RayTracing, siddon.cu.cc:293-331).  KIND 0: gather (returns
// the projection), 1: scatter seg*scale, 2: count non-zero segments.  The
// per-axis state lives in scalar registers (no runtime-indexed arrays).
// Value of the segment (last_t, nt] of a ray, the reference's per-voxel
// factor of RayTracing / BackRayTracing (siddon.cu.cc:303-322): l / L^2,
// times the TOF Gaussian when TOF is on.  Returns false when the TOF test
// (t^2 / sigma^2 < 9) drops the segment.  Every Siddon walk forms its values
// here, so the untiled and the tiled walks produce the same bits.
template <bool TOF>
__device__ __forceinline__ bool seg_value(const SidParams &p, float weight, double t_tof,
                                          double nt, double last_t, float &g)
{
    const double l_d = __dsub_rn(nt, last_t);
    const float l_len = (float)l_d;
    if (TOF) {
        const double bs2 = (double)p.tof_bin * (double)p.tof_bin;
        double t = __dsub_rn(t_tof, __dmul_rn(0.5, __dadd_rn(nt, last_t)));
        double sigma2 = __dadd_rn((double)p.tof_sigma2,
                                  __ddiv_rn(__dadd_rn(__dmul_rn(l_d, l_d), bs2), 12.0));
        double tt = __ddiv_rn(__dmul_rn(t, t), sigma2);
        if (!(tt < 9.0))
            return false;
        // bin * exp(-t2/2) / sqrt(2 pi sigma2)
        g = weight * l_len * p.tof_bin * __expf(-0.5f * (float)tt) *
            __frsqrt_rn(6.283185307179586f * (float)sigma2);
        return true;
    }
    g = weight * l_len;
    return true;
}

// Walk the ray (RayTracing, siddon.cu.cc:293-331).  KIND 0: gather (returns
// the projection), 1: scatter seg*scale, 2: count non-zero segments.  The
// per-axis state lives in scalar registers (no runtime-indexed arrays).
template <int KIND, bool TOF>
__device__ __forceinline__ float walk(const SidParams &p, const Walk &w, float scale, int *count)
{
    int off = w.pos[0] * p.s[0] + w.pos[1] * p.s[1] + w.pos[2] * p.s[2];
    int ooff = w.pos[0] * p.o[0] + w.pos[1] * p.o[1] + w.pos[2] * p.o[2];
    const int ds0 = w.step[0] * p.s[0], ds1 = w.step[1] * p.s[1], ds2 = w.step[2] * p.s[2];
    const int os0 = w.step[0] * p.o[0], os1 = w.step[1] * p.o[1], os2 = w.step[2] * p.o[2];
    double n0 = w.nextT[0], n1 = w.nextT[1], n2 = w.nextT[2];
    const double t0 = w.deltaT[0], t1 = w.deltaT[1], t2 = w.deltaT[2];
    int b0 = w.inBuf[0], b1 = w.inBuf[1], b2 = w.inBuf[2];
    double last_t = 0.0;
    float pv = 0.0f;
    int stillin = b0 * b1 * b2;
    int cap = p.g[0] + p.g[1] + p.g[2] + 3;
    while (stillin && cap-- > 0) {
        int i = (n0 <= n1) ? ((n0 <= n2) ? 0 : 2) : ((n1 <= n2) ? 1 : 2);
        double nt = i == 0 ? n0 : (i == 1 ? n1 : n2);
        float g;
        if (seg_value<TOF>(p, w.weight, w.t_tof, nt, last_t, g)) {
            if (KIND == 2)
                *count += (__dsub_rn(nt, last_t) > 0.0);
            else if (KIND == 0)
                pv += p.img[off] * g;
            else
                unsafeAtomicAdd(p.out + ooff, g * scale);
        }
        last_t = nt;
        if (i == 0) {
            n0 = __dadd_rn(nt, t0);
            off += ds0;
            ooff += os0;
            stillin = --b0;
        } else if (i == 1) {
            n1 = __dadd_rn(nt, t1);
            off += ds1;
            ooff += os1;
            stillin = --b1;
        } else {
            n2 = __dadd_rn(nt, t2);
            off += ds2;
            ooff += os2;
            stillin = --b2;
        }
    }
    return pv;
}


} // namespace
} // namespace srfx
