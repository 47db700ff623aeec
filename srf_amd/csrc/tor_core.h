// Device-side TOF tube-of-response (TOR) projector shared by the HIP
// translation units (tor.hip: single-job entry points; batch.hip: batched
// persistent launch).  Each translation unit gets its own internal copy.
//
// Semantics follow the reference CUDA op src/cpp/tf_custom_ops/tor.cu.cc (and
// tor.cc for the op boundary).  The reference launches one 32x1024 kernel per
// image slice and re-tests every event against every slice (tor.cu.cc:377-391).
// Here one lane owns one event for the whole call: it finds the event's
// strictly crossed slices from its end points, walks only those, and keeps
// the running projection in a register.
//
// Precision.  The tube is narrow (sigma = kernel_width/6, 0.67 mm at
// kernel_width 4) and the crossing point sits hundreds of mm from the origin,
// so FP32 rounding of the crossing point alone moves the weights by ~1e-5.
// Positions are therefore carried in FP64 (full rate on CDNA): slice centre,
// crossing point, patch origin and the offset of the crossing point from the
// patch's first voxel centre.  In-patch offsets, squared distances and weights
// are FP32.  Every include/exclude decision is the FP64 one of the restated
// reference: the FP32 cut-off test is re-evaluated in FP64 whenever d^2 lies
// within 1e-4 mm^2 of the cut-off, and a patch origin within 1e-9 voxel of a
// voxel boundary is recomputed with the reference's exact divisions.
#pragma once

#include <math.h>

#include "common.h"

namespace srfx {
namespace {

enum TorMode { kProj = 0, kBack = 1, kMap = 2, kAdj = 3, kCount = 4, kOsem = 5 };

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kThrMargin = 1.0e-4f; // mm^2 band around the cut-off decided in FP64

} // namespace

struct TorParams {
    const float *__restrict__ lors;
    int64_t n, ld;
    const float *__restrict__ val; // lors_value (bp, map) or g (adjoint)
    const float *__restrict__ img; // image read by proj / osem
    float *out;                    // proj[n] (proj) or image (bp, map, adj, osem)
    int out_step;                  // proj: element step of out (1, or 12 into event records)
    int32_t *nnz;                  // count: per-event counts (nullable)
    unsigned long long *nnz_total; // count: running total (nullable)
    int gx, gy, gz;
    int sx, sy, sz; // strides of the read image
    int img_bytes;  // bytes of a read image in contiguous op-frame layout (frame gathers)
    int ox, oy, oz; // strides of the written image
    double inter_x, inter_y, inter_z;
    double rinter_x, rinter_y, rinter_z;
    double l_bound, b_bound;
    double dx_half, dy_half; // (half - 0.5) * inter_x, (half - 0.5) * inter_y
    float fdx_half, fdy_half; // the same in FP32
    double zbase; // center_z - (lz - inter_z)/2: slice s is at zbase + s*inter_z
    double tor_sigma2;
    float finter_x, finter_y;
    float rfinter_x, rfinter_y;
    int patch, half;
    int box;        // compiled candidate box (3, 4) or 0 (chord walk only)
    int full_patch; // evaluate every patch voxel (reference loop) instead of the chords
    int geo;
    int use_tof;
    float tse_inv; // 1 / (tof_sigma2 + tof_bin^2/12)
    float tofk;    // tof_bin / sqrt(2 pi tof_sigma2)
};


// An event as the walks carry it.  The FP64 geometry that only the rare
// exact decisions need is not kept in registers: e1 / e2 locate the event's
// FP32 end points (x1, y1, z1 at e1[0], e1[ld], e1[2 ld]; x2, y2, z2 and the
// sigma2 factor at e2[0], e2[ld], e2[2 ld], e2[fo ld]) and exact_geom forms
// it again from them.
struct Event {
    const float *e1, *e2;
    int64_t ld;
    int fo;
    float ux, uy;     // FP32 direction cosines (the exact ones are recomputed where needed)
    float thr;        // FP32 image of 9 sigma_c^2
    float cw;         // per-voxel exponent factor (log2 domain)
    float ctof;       // per-slice TOF exponent factor (log2 domain)
    int s_lo, s_hi;   // strictly crossed slices
};

// The reference's FP64 event geometry (tor.cu.cc:22-48): end point 1, the
// differences d = p2 - p1 of the FP32 end points, |d|^2 and sigma_c^2.
struct ExactGeom {
    double x1, y1, z1, d0, d1, d2;
    double dis2; // |d|^2
    double sc_d; // sigma_c^2 = tor_sigma2 (* sigma2 factor)
};

namespace {

__device__ __forceinline__ double slice_z(const TorParams &p, int s)
{
    return __dadd_rn(p.zbase, __dmul_rn((double)s, p.inter_z));
}

__device__ __forceinline__ ExactGeom exact_geom(const TorParams &p, float x1, float y1, float z1,
                                                float x2, float y2, float z2, float f)
{
    ExactGeom g;
    g.x1 = x1;
    g.y1 = y1;
    g.z1 = z1;
    g.d0 = (double)x2 - (double)x1;
    g.d1 = (double)y2 - (double)y1;
    g.d2 = (double)z2 - (double)z1;
    g.dis2 = __dadd_rn(__dadd_rn(__dmul_rn(g.d0, g.d0), __dmul_rn(g.d1, g.d1)),
                       __dmul_rn(g.d2, g.d2));
    g.sc_d = p.geo ? p.tor_sigma2 * (double)f : p.tor_sigma2;
    return g;
}

// The geometry again from the event's end points (the rare exact paths).
__device__ __forceinline__ ExactGeom exact_geom(const TorParams &p, const Event &ev)
{
    const int64_t ld = ev.ld;
    return exact_geom(p, ev.e1[0], ev.e1[ld], ev.e1[2 * ld], ev.e2[0], ev.e2[ld], ev.e2[2 * ld],
                      ev.e2[ev.fo * ld]);
}

// Slices s with 0 < (z_s - z1)/d2 < 1 (tor.cu.cc:29-33), i.e. z_s strictly
// between z1 and z2.  Returns false when there are none.
__device__ __forceinline__ bool slice_range(const TorParams &p, const ExactGeom &ev, int &s_lo,
                                            int &s_hi)
{
    double za, zb; // open interval (za, zb)
    if (ev.d2 > 0.0) {
        za = ev.z1;
        zb = ev.z1 + ev.d2;
    } else if (ev.d2 < 0.0) {
        za = ev.z1 + ev.d2;
        zb = ev.z1;
    } else {
        return false;
    }
    if (!(za > -1.0e30 && zb < 1.0e30))
        return false;
    // starting guesses (any rounding is corrected by the exact compares below)
    double fa = (za - p.zbase) * p.rinter_z, fb = (zb - p.zbase) * p.rinter_z;
    if (fb < -1.0 || fa > (double)p.gz + 1.0)
        return false;
    int a = (int)fmax(floor(fa), -1.0);
    int b = (int)fmin(ceil(fb), (double)p.gz);
    // first slice strictly above za, last strictly below zb (exact compares)
#pragma clang loop vectorize(disable) unroll(disable)
    while (a > 0 && slice_z(p, a - 1) > za)
        a--;
#pragma clang loop vectorize(disable) unroll(disable)
    while (a <= b && !(slice_z(p, a) > za))
        a++;
#pragma clang loop vectorize(disable) unroll(disable)
    while (b < p.gz - 1 && slice_z(p, b + 1) < zb)
        b++;
#pragma clang loop vectorize(disable) unroll(disable)
    while (b >= a && !(slice_z(p, b) < zb))
        b--;
    if (ev.d2 > 0.0) {
        // ratio < 1 is tested on the rounded quotient: drop a slice whose
        // quotient rounds to 1.
#pragma clang loop vectorize(disable) unroll(disable)
        while (b >= a && !(((slice_z(p, b) - ev.z1) / ev.d2) < 1.0))
            b--;
    } else {
#pragma clang loop vectorize(disable) unroll(disable)
        while (b >= a && !(((slice_z(p, a) - ev.z1) / ev.d2) < 1.0))
            a++;
    }
    s_lo = a < 0 ? 0 : a;
    s_hi = b > p.gz - 1 ? p.gz - 1 : b;
    return s_lo <= s_hi;
}

// Raw event record: end points, TOF centre (zero for Maplors rows) and the
// sigma2 factor.
} // namespace

struct RawEvent {
    float x1, y1, z1, x2, y2, z2, xc, yc, zc, f;
};

namespace {

template <int MODE>
__device__ __forceinline__ void fetch_event(const TorParams &p, int64_t e, RawEvent &r)
{
    const float *L = p.lors;
    const int64_t ld = p.ld;
    r.x1 = __builtin_nontemporal_load(L + 0 * ld + e);
    r.y1 = __builtin_nontemporal_load(L + 1 * ld + e);
    r.z1 = __builtin_nontemporal_load(L + 2 * ld + e);
    r.x2 = __builtin_nontemporal_load(L + 3 * ld + e);
    r.y2 = __builtin_nontemporal_load(L + 4 * ld + e);
    r.z2 = __builtin_nontemporal_load(L + 5 * ld + e);
    if (MODE == kMap) {
        r.f = __builtin_nontemporal_load(L + 6 * ld + e);
        r.xc = r.yc = r.zc = 0.0f;
    } else {
        r.xc = __builtin_nontemporal_load(L + 6 * ld + e);
        r.yc = __builtin_nontemporal_load(L + 7 * ld + e);
        r.zc = __builtin_nontemporal_load(L + 8 * ld + e);
        r.f = __builtin_nontemporal_load(L + 9 * ld + e);
    }
}

// Per-event set-up (tor.cu.cc:18-40 hoisted out of the slice loop) from the
// event's FP32 row r and its FP64 geometry g; the caller sets ev.e1 / e2.
template <int MODE>
__device__ __forceinline__ bool init_event(const TorParams &p, const RawEvent &r,
                                           const ExactGeom &g, Event &ev)
{
    const double dis = __dsqrt_rn(g.dis2);
    // FP32 direction cosines for the weights and the margin-guarded cut-off
    // test; the exact d0 / dis of the reference is recomputed by the FP64
    // decision (inside_exact).
    const double rdis = 1.0 / dis;
    ev.ux = (float)(g.d0 * rdis);
    ev.uy = (float)(g.d1 * rdis);
    ev.thr = (float)(9.0 * g.sc_d);
    float inv_sc = (float)(1.0 / g.sc_d);
    if (MODE != kMap && MODE != kCount && p.use_tof) {
        ev.cw = kLog2e * 0.5f * (p.tse_inv - inv_sc);
        ev.ctof = -0.5f * kLog2e * p.tse_inv;
    } else {
        ev.cw = -0.5f * kLog2e * inv_sc;
        ev.ctof = 0.0f;
    }
    ev.s_lo = 1;
    ev.s_hi = 0;
    if (!(g.sc_d > 0.0) || !(dis > 0.0) || !(dis < 1.0e30))
        return false;
    // (d1, d2) = 0 means no crossed slice (d2 = 0); keeps rden finite
    if (!(g.d1 * g.d1 + g.d2 * g.d2 > 0.0))
        return false;
    return slice_range(p, g, ev.s_lo, ev.s_hi);
}

__device__ __forceinline__ ExactGeom exact_geom(const TorParams &p, const RawEvent &r)
{
    return exact_geom(p, r.x1, r.y1, r.z1, r.x2, r.y2, r.z2, r.f);
}

// Per-event geometry carried across slices.  The crossing point of slice s,
// in voxel units from the grid corner, is affine in s: u(s) = ax + s*bx,
// w(s) = ay + s*by (FP64).  The vector from the crossing point to the TOF
// centre is v0 - (s - s_lo)*dv (FP32).
} // namespace

struct Track {
    double ax, bx, ay, by;
    // TOF exponent of slice s (log2 domain): tc0 + s (tc1 + s tc2), the
    // squared distance from the TOF centre to the crossing point times ctof
    // expanded in s (FP64, so the expansion loses nothing)
    double tc0, tc1, tc2;
    // in-slice ellipse d2 < tm: d2 = a (dx - cxy dy)^2 + q dy^2
    float A;     // 1 / (1 - ux^2)
    float cxy;   // ux uy / (1 - ux^2)
    float q;     // uz^2 / (1 - ux^2)
    float tm;    // cut-off padded so the candidate set contains every kept voxel
    float dymax; // sqrt(tm / q): half height of the ellipse
    float hx;    // sqrt(tm (A + cxy^2 / q)): half width of the ellipse
    int nx, ny;  // voxel columns / rows any slice's ellipse can span
};

namespace {

// The track of an event with crossed slices, from its FP64 geometry g and
// its row r (TOF centre).
__device__ __forceinline__ void make_track(const TorParams &p, const Event &ev,
                                           const ExactGeom &g, const RawEvent &r, Track &t)
{
    const double rd2 = 1.0 / g.d2;
    double r0 = (p.zbase - g.z1) * rd2;        // ratio at slice 0
    double rs = p.inter_z * rd2;               // ratio step per slice
    t.ax = (g.x1 + g.d0 * r0 - p.l_bound) * p.rinter_x;
    t.bx = g.d0 * rs * p.rinter_x;
    t.ay = (g.y1 + g.d1 * r0 - p.b_bound) * p.rinter_y;
    t.by = g.d1 * rs * p.rinter_y;
    // TOF centre minus the crossing point of slice 0, and its step per slice
    const double vx = (double)r.xc - (g.x1 + g.d0 * r0), vy = (double)r.yc - (g.y1 + g.d1 * r0);
    const double vz = (double)r.zc - slice_z(p, 0);
    const double dx = g.d0 * rs, dy = g.d1 * rs, dz = p.inter_z;
    const double ct = (double)ev.ctof;
    t.tc0 = ct * (vx * vx + vy * vy + vz * vz);
    t.tc1 = -2.0 * ct * (vx * dx + vy * dy + vz * dz);
    t.tc2 = ct * (dx * dx + dy * dy + dz * dz);
    // a = 1 - ux^2 = (d1^2 + d2^2) / |d|^2, formed without cancellation
    // (d2 != 0 for an event with crossed slices)
    const double den = g.d1 * g.d1 + g.d2 * g.d2;
    const double rden = 1.0 / den;
    t.A = (float)(g.dis2 * rden);
    t.cxy = (float)(g.d0 * g.d1 * rden);
    t.q = (float)(g.d2 * g.d2 * rden);
    t.tm = ev.thr * (1.0f + 1.0e-5f) + 2.0f * kThrMargin;
    t.dymax = sqrtf(t.tm / t.q);
    t.hx = sqrtf(t.tm * (t.A + t.cxy * t.cxy / t.q));
    // integers in an open interval of length 2h/pitch, with the index
    // rounding allowance of the box start (1e-4) on both ends
    t.nx = (int)fminf(floorf(2.0f * t.hx * p.rfinter_x + 2.0e-4f), 1.0e6f) + 1;
    t.ny = (int)fminf(floorf(2.0f * t.dymax * p.rfinter_y + 2.0e-4f), 1.0e6f) + 1;
}

// Event and track of row e of the raw LOR rows (MODE kMap: 7 columns, else
// 10).  Returns false for an event without crossed slices (t unset).
template <int MODE>
__device__ __forceinline__ bool load_event(const TorParams &p, int64_t e, Event &ev, Track &t)
{
    RawEvent r;
    fetch_event<MODE>(p, e, r);
    const ExactGeom g = exact_geom(p, r);
    ev.e1 = p.lors + e;
    ev.e2 = p.lors + 3 * p.ld + e;
    ev.ld = p.ld;
    ev.fo = MODE == kMap ? 3 : 6;
    if (!init_event<MODE>(p, r, g, ev))
        return false;
    make_track(p, ev, g, r, t);
    return true;
}

} // namespace

struct Cross {
    int ix0, iy0;
    float dx0, dy0; // crossing point minus centre of the first patch voxel, mm
};

namespace {
// Walk record of one event: everything a walk needs, produced once by
// init_event and make_track (a plan's count pass) and unpacked by each of the
// event's tile listings instead of recomputing it.  A record array of n
// events holds n heads of 8 float4 (128 bytes, one cache line: all that the
// common path of a walk reads) and then n tails of 1 float4:
//   head [ax bx] [ay by] [tc0 tc1] (FP64 pairs) [tc2 (FP64) dymax hx]
//        [ux uy thr cw] [ctof s_lo s_hi nx|ny<<16] [A cxy q tm] [x1 y1 z1 value]
//   tail [x2 y2 z2 f]
// value is the scale of the event's backprojection in a plan's OSEM step
// (1/p).  An event without a crossed slice (or otherwise invalid) has
// s_lo = 1, s_hi = 0.
constexpr int kWalkF4 = 9;    // float4 per event
constexpr int kWalkHead = 8;  // float4 per head
constexpr int kWalkValue = 31; // float index of the value in a head

__device__ __forceinline__ float4 dpair(double a, double b)
{
    return make_float4(__int_as_float(__double2loint(a)), __int_as_float(__double2hiint(a)),
                       __int_as_float(__double2loint(b)), __int_as_float(__double2hiint(b)));
}

__device__ __forceinline__ void undpair(const float4 &f, double &a, double &b)
{
    a = __hiloint2double(__float_as_int(f.y), __float_as_int(f.x));
    b = __hiloint2double(__float_as_int(f.w), __float_as_int(f.z));
}

__device__ __forceinline__ void pack_walk(const TorParams &p, const RawEvent &r, const Event &ev,
                                          const Track &t, bool valid, float4 *rec, int64_t e)
{
    float4 *w = rec + kWalkHead * e;
    w[0] = dpair(t.ax, t.bx);
    w[1] = dpair(t.ay, t.by);
    w[2] = dpair(t.tc0, t.tc1);
    w[3] = dpair(t.tc2, 0.0);
    w[3].z = t.dymax;
    w[3].w = t.hx;
    w[4] = make_float4(ev.ux, ev.uy, ev.thr, ev.cw);
    w[5] = make_float4(ev.ctof, __int_as_float(valid ? ev.s_lo : 1),
                       __int_as_float(valid ? ev.s_hi : 0),
                       __int_as_float((t.nx & 0xffff) | (t.ny << 16)));
    w[6] = make_float4(t.A, t.cxy, t.q, t.tm);
    w[7] = make_float4(r.x1, r.y1, r.z1, 0.0f);
    rec[kWalkHead * p.n + e] = make_float4(r.x2, r.y2, r.z2, r.f);
}

// Inverse of pack_walk for event e of a record array of p.n events: the Event
// and Track fields the walks read; the end points stay in the record for the
// rare exact decisions.  Returns the record's value.
__device__ __forceinline__ float unpack_walk(const TorParams &p, const float4 *__restrict__ rec,
                                             int64_t e, Event &ev, Track &t)
{
    const float4 *w = rec + kWalkHead * e;
    const float4 a = w[0], b = w[1], c = w[2], d = w[3], f = w[4], g = w[5], h = w[6], k = w[7];
    undpair(a, t.ax, t.bx);
    undpair(b, t.ay, t.by);
    undpair(c, t.tc0, t.tc1);
    t.tc2 = __hiloint2double(__float_as_int(d.y), __float_as_int(d.x));
    t.dymax = d.z;
    t.hx = d.w;
    ev.ux = f.x;
    ev.uy = f.y;
    ev.thr = f.z;
    ev.cw = f.w;
    ev.ctof = g.x;
    ev.s_lo = __float_as_int(g.y);
    ev.s_hi = __float_as_int(g.z);
    const int nxy = __float_as_int(g.w);
    t.nx = nxy & 0xffff;
    t.ny = nxy >> 16;
    t.A = h.x;
    t.cxy = h.y;
    t.q = h.z;
    t.tm = h.w;
    ev.e1 = reinterpret_cast<const float *>(w + 7);
    ev.e2 = reinterpret_cast<const float *>(rec + kWalkHead * p.n + e);
    ev.ld = 1;
    ev.fo = 3;
    return k.w;
}

} // namespace

namespace {

// The reference's crossing point of slice s (tor.cu.cc:30-36) in FP64 with
// its operation order.
__device__ __forceinline__ void cross_point_exact(const TorParams &p, const ExactGeom &ev, int s,
                                                  double &cx, double &cy)
{
    double ratio = __ddiv_rn(__dsub_rn(slice_z(p, s), ev.z1), ev.d2);
    cx = __dadd_rn(ev.x1, __dmul_rn(ev.d0, ratio));
    cy = __dadd_rn(ev.y1, __dmul_rn(ev.d1, ratio));
}

__device__ __forceinline__ double mesh_x(const TorParams &p, int ix)
{
    return __dadd_rn(__dmul_rn(p.inter_x, (double)ix + 0.5), p.l_bound);
}
__device__ __forceinline__ double mesh_y(const TorParams &p, int iy)
{
    return __dadd_rn(__dmul_rn(p.inter_y, (double)iy + 0.5), p.b_bound);
}

// Patch origin (tor.cu.cc:135-136) and offsets for slice s.  Near a voxel
// boundary the origin is recomputed with the reference's exact divisions.
__device__ __forceinline__ void cross_at(const TorParams &p, const Event &ev, const Track &t, int s,
                                         Cross &c)
{
    double u = t.ax + (double)s * t.bx;
    double w = t.ay + (double)s * t.by;
    // trunc as the reference's (int) cast; the fraction u - trunc(u) lies in
    // (-1, 1), here rounded to FP32 (relative error 2^-24).  u near an
    // integer means a fraction near 0 or +-1: every |fraction| below 1e-9
    // or above 1 - 1e-9 is caught as |fu| < 1e-9 or |fu| >= 1 - 2^-24 (a
    // fraction above 1 - 1e-9 rounds to 1), far wider than the FP64 error of
    // u (~1e-13 voxel).  A non-short-circuit or, so one branch.
    const int iu = (int)u, iw = (int)w;
    const float fu = (float)(u - (double)iu), fw = (float)(w - (double)iw);
    const float alo = fminf(fabsf(fu), fabsf(fw)), ahi = fmaxf(fabsf(fu), fabsf(fw));
    if ((int)(alo < 1.0e-9f) | (int)(ahi >= 1.0f - 0x1p-24f)) {
        double cx, cy;
        cross_point_exact(p, exact_geom(p, ev), s, cx, cy);
        c.ix0 = (int)__ddiv_rn(__dsub_rn(cx, p.l_bound), p.inter_x) - p.half;
        c.iy0 = (int)__ddiv_rn(__dsub_rn(cy, p.b_bound), p.inter_y) - p.half;
        c.dx0 = (float)(cx - mesh_x(p, c.ix0));
        c.dy0 = (float)(cy - mesh_y(p, c.iy0));
    } else {
        c.ix0 = iu - p.half;
        c.iy0 = iw - p.half;
        // (u - (ix0 + 0.5)) * inter = fu * inter + (half - 0.5) * inter, in
        // FP32 from the rounded fraction: error ~3e-7 mm, whose effect on
        // d2 stays far inside kThrMargin
        c.dx0 = fmaf(fu, p.finter_x, p.fdx_half);
        c.dy0 = fmaf(fw, p.finter_y, p.fdy_half);
    }
}

// The value, unchanged, but opaque to the optimizer: computations that depend
// on it cannot be hoisted out of the (rarely taken) branch that contains them.
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Exact FP64 cut-off decision of the reference (tor.cu.cc:57-71) for voxel
// (ix, iy) of slice s.
__device__ __forceinline__ bool inside_exact(const TorParams &p, const Event &ev, int s, int ix,
                                             int iy)
{
    s = opaque(s);
    const ExactGeom g = exact_geom(p, ev);
    double cx, cy;
    cross_point_exact(p, g, s, cx, cy);
    double dis = __dsqrt_rn(g.dis2);
    double ux = __ddiv_rn(g.d0, dis), uy = __ddiv_rn(g.d1, dis);
    double dx = __dsub_rn(cx, mesh_x(p, ix)), dy = __dsub_rn(cy, mesh_y(p, iy));
    double rc = __dadd_rn(__dmul_rn(dx, ux), __dmul_rn(dy, uy));
    double d2 = __dsub_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(rc, rc));
    return d2 < 9.0 * g.sc_d;
}

// Per-slice TOF factor tof_bin/sqrt(2 pi s2) * exp(-D/(2 tse)), D the squared
// distance from the TOF centre to the crossing point; the reference's
// d2_tof = D - d2 (tor.cu.cc:87) has its -d2 part folded into the per-voxel
// exponent.  D / (2 tse) in the log2 domain is the track's quadratic in s,
// evaluated in FP64 and rounded once.
__device__ __forceinline__ float tof_exp(const TorParams &p, const Event &ev, const Track &t,
                                         int s)
{
    const double sd = (double)s;
    return __builtin_amdgcn_exp2f((float)fma(fma(t.tc2, sd, t.tc1), sd, t.tc0));
}

__device__ __forceinline__ float tof_factor(const TorParams &p, const Event &ev, const Track &t,
                                            int s)
{
    return p.tofk * tof_exp(p, ev, t, s);
}

// Offset of voxel column i (row j) from the crossing point, and dt = d2 - thr,
// d2 = |D|^2 - (D.u)^2 the squared distance of a voxel centre from the LOR
// axis, with the cut-off folded into the row term.  Every walk evaluates them
// through these helpers with contraction off and the fused steps written out,
// so that the box walk, the chord walk and the full patch produce the same
// bits for the same voxel.  For a voxel near the cut-off the FP32 error of dt
// is far below kThrMargin: |D|^2 <= 3 d2 for an event along its dominant
// axis, so every term is at most a few times thr.
__device__ __forceinline__ float voxel_offset(float o0, int i, float step)
{
    return fmaf(-(float)i, step, o0);
}

// Column terms (dx^2, ux dx) and row terms (dy^2 - thr, uy dy).
__device__ __forceinline__ void col_terms(float dx, float ux, float &dx2, float &uxdx)
{
#pragma clang fp contract(off)
    dx2 = dx * dx;
    uxdx = ux * dx;
}

__device__ __forceinline__ void row_terms(float dy, float uy, float thr, float &dyt, float &uydy)
{
#pragma clang fp contract(off)
    dyt = fmaf(dy, dy, -thr);
    uydy = uy * dy;
}

__device__ __forceinline__ float tube_dt(float dx2, float uxdx, float dyt, float uydy)
{
#pragma clang fp contract(off)
    float rc = uxdx + uydy;
    return fmaf(-rc, rc, dx2 + dyt);
}

__device__ __forceinline__ float voxel_dt(const Event &ev, float dx, float dy)
{
    float dx2, uxdx, dyt, uydy;
    col_terms(dx, ev.ux, dx2, uxdx);
    row_terms(dy, ev.uy, ev.thr, dyt, uydy);
    return tube_dt(dx2, uxdx, dyt, uydy);
}

// Weight of a kept voxel: amp exp(-d2 / (2 sigma^2)), the exponent as
// dt cw + thr cw.
__device__ __forceinline__ float voxel_weight(const Event &ev, float amp, float dt)
{
    return amp * __builtin_amdgcn_exp2f(fmaf(dt, ev.cw, ev.thr * ev.cw));
}

// One voxel: exact include test, then gather / scatter / count.
template <int KIND>
__device__ __forceinline__ void visit(const TorParams &p, const Event &ev, int s, int ix, int iy,
                                      float dx, float dy, float amp, float scale, float &acc)
{
    const float dt = voxel_dt(ev, dx, dy);
    bool in = dt < -kThrMargin;
    if (!in && dt <= kThrMargin)
        in = inside_exact(p, ev, s, ix, iy);
    if (in) {
        if (KIND == 2) {
            acc += 1.0f;
        } else {
            float w = voxel_weight(ev, amp, dt);
            if (KIND == 0)
                acc = fmaf(w, p.img[ix * p.sx + iy * p.sy + s * p.sz], acc);
            else
                unsafeAtomicAdd(p.out + (ix * p.ox + iy * p.oy + s * p.oz), w);
        }
    }
}

// Walk one slice's patch (tor.cu.cc:124-166 and siblings), j outer, i inner,
// voxels outside the grid skipped.  FULL evaluates all P x P voxels as the
// reference loop does; otherwise each row only evaluates the voxels whose
// centres fall in the analytically computed chord of the (padded) tube
// ellipse, a superset of the kept voxels, and the exact test decides.
//   KIND 0: returns sum w*img; 1: out += w*scale; 2: returns the count.
template <int KIND, bool FULL>
__device__ __forceinline__ float patch_walk(const TorParams &p, const Event &ev, const Track &t,
                                            const Cross &c, float amp, int s, float scale)
{
    const int P = p.patch;
    // A patch entirely outside the grid contributes nothing; this also keeps
    // the loop bounds below free of overflow for saturated origins.
    if (c.ix0 <= -P || c.ix0 >= p.gx || c.iy0 <= -P || c.iy0 >= p.gy)
        return 0.0f;
    int i_lo = c.ix0 < 0 ? -c.ix0 : 0;
    int i_hi = min(P, p.gx - c.ix0) - 1;
    int j_lo = c.iy0 < 0 ? -c.iy0 : 0;
    int j_hi = min(P, p.gy - c.iy0) - 1;
    float acc = 0.0f;
    if (!FULL) {
        const float eps = 1.0e-4f;
        j_lo = max(j_lo, (int)ceilf((c.dy0 - t.dymax) * p.rfinter_y - eps));
        j_hi = min(j_hi, (int)floorf((c.dy0 + t.dymax) * p.rfinter_y + eps));
    }
    for (int j = j_lo; j <= j_hi; j++) {
        int iy = c.iy0 + j;
        float dy = voxel_offset(c.dy0, j, p.finter_y);
        int il = i_lo, ih = i_hi;
        if (!FULL) {
            float tt = t.tm - dy * dy * t.q;
            if (tt <= 0.0f)
                continue;
            float hw = __builtin_amdgcn_sqrtf(tt * t.A);
            float ctr = c.dx0 - t.cxy * dy;
            const float eps = 1.0e-4f;
            il = max(il, (int)ceilf((ctr - hw) * p.rfinter_x - eps));
            ih = min(ih, (int)floorf((ctr + hw) * p.rfinter_x + eps));
        }
        for (int i = il; i <= ih; i++) {
            float dx = voxel_offset(c.dx0, i, p.finter_x);
            visit<KIND>(p, ev, s, c.ix0 + i, iy, dx, dy, amp, scale, acc);
        }
    }
    return acc;
}

// Index of patch voxel (i, j) of one slice in the accessed array: the
// strided image, or an LDS tile.
} // namespace

struct PatchIdx {
    int base, si, sj;
    __device__ __forceinline__ int operator()(int i, int j) const { return base + i * si + j * sj; }
};

namespace {

// One slice through the event's candidate box: the NB x NB voxels (at most
// t.nx columns, t.ny rows) starting at the first column / row whose centre can
// lie inside the padded tube ellipse, clipped to patch voxels [i_lo, i_hi] x
// [j_lo, j_hi] (patch, grid and tile bounds).  Every kept voxel of the slice
// is in the box when t.nx, t.ny <= NB (the caller checks).  Three passes:
// all squared distances and the FP32 include decisions; the FP64 decision for
// the rare voxels within the margin of the cut-off; the weights of the kept
// voxels.  The gather keeps the decisions as per-voxel lane masks and sums in
// the reference's order (rows outer, columns inner) with the arithmetic of
// patch_walk; the scatter and the count decide from d2 - thr directly.
//   KIND 0: returns sum w * src[idx]; 1: dst[idx] += w, amp carrying the
//   scale (LDS: LDS atomics in dst's type, else FP32 global atomics);
//   2: returns the count.
//   LDS: 0 strided image; 1 LDS tile with unit column stride (idx.si = 1);
//   2 LDS tile with unit row stride (idx.sj = 1).
template <int KIND, int BX, int BY, int LDS, typename D>
__device__ __forceinline__ float slice_box(const TorParams &p, const Event &ev, const Track &t,
                                           const Cross &c, float amp, int s, float scale, int i_lo,
                                           int i_hi, int j_lo, int j_hi, const PatchIdx &idx,
                                           const float *src, D *dst)
{
    constexpr int NC = BX * BY;
    const float eps = 1.0e-4f;
    const int ib0 = (int)ceilf((c.dx0 - t.hx) * p.rfinter_x - eps);
    const int jb0 = (int)ceilf((c.dy0 - t.dymax) * p.rfinter_y - eps);
    const int ib = max(i_lo, ib0), ie = min(i_hi, ib0 + t.nx - 1);
    const int jb = max(j_lo, jb0), je = min(j_hi, jb0 + t.ny - 1);
    // LDS tiles have unit column (LDS 1) or row (LDS 2) stride: one address
    // per box row (column), formed after the decisions (opaque, so not
    // re-formed per voxel inside the branches), and the other index as an
    // immediate offset
    int row[BY], col[BX];
    // dt = d2 - thr of every candidate from its column and row terms; the
    // smallest |dt| decides whether the FP64 decision is needed at all
    float cx2[BX], cux[BX], rt2[BY], ruy[BY];
    if constexpr (KIND == 0) {
        // offsets as the patch walks form them (same bits)
#pragma unroll
        for (int k = 0; k < BX; k++)
            col_terms(voxel_offset(c.dx0, ib + k, p.finter_x), ev.ux, cx2[k], cux[k]);
#pragma unroll
        for (int k = 0; k < BY; k++)
            row_terms(voxel_offset(c.dy0, jb + k, p.finter_y), ev.uy, ev.thr, rt2[k], ruy[k]);
        // Gather.  Kept: dt < -margin; within |dt| <= margin the FP64
        // decision.  The decisions stay per-candidate lane masks (measured
        // faster than a packed mask for the gather); the kept voxels are
        // summed in the reference's order (rows outer, columns inner), all
        // loads first.
        float dt[NC];
        float gap = INFINITY; // the smallest |dt| over the box
#pragma unroll
        for (int jj = 0; jj < BY; jj++) {
#pragma unroll
            for (int ii = 0; ii < BX; ii++) {
                const int k = jj * BX + ii;
                dt[k] = tube_dt(cx2[ii], cux[ii], rt2[jj], ruy[jj]);
                gap = fminf(gap, fabsf(dt[k]));
            }
        }
        bool in[NC];
#pragma unroll
        for (int k = 0; k < NC; k++)
            in[k] = (ib + k % BX <= ie) & (jb + k / BX <= je) & (dt[k] < -kThrMargin);
        if (gap <= kThrMargin) {
            // rare: the reference's FP64 decision, one voxel at a time
            uint32_t near = 0, add = 0;
#pragma unroll
            for (int k = 0; k < NC; k++) {
                const bool ok = (ib + k % BX <= ie) & (jb + k / BX <= je);
                near |= (uint32_t)(ok & (fabsf(dt[k]) <= kThrMargin)) << k;
            }
            for (uint32_t m = (uint32_t)opaque((int)near); m; m &= m - 1) {
                const int k = __builtin_ctz(m);
                if (inside_exact(p, ev, s, c.ix0 + ib + k % BX, c.iy0 + jb + k / BX))
                    add |= 1u << k;
            }
#pragma unroll
            for (int k = 0; k < NC; k++)
                in[k] |= ((add >> k) & 1u) != 0;
        }
        static_assert(LDS != 2, "gathers read x-fastest tiles");
#pragma unroll
        for (int jj = 0; jj < BY; jj++)
            row[jj] = LDS ? opaque(idx(ib, jb + jj)) : 0;
        float v[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int a = LDS ? row[k / BX] + k % BX : idx(ib + k % BX, jb + k / BX);
            v[k] = in[k] ? src[a] : 0.0f;
        }
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < NC; k++)
            if (in[k])
                acc = fmaf(voxel_weight(ev, amp, dt[k]), v[k], acc);
        return acc;
    } else {
        // Scatter and count: no packed decisions.  Columns and rows outside
        // the clip get an infinite term, so that their candidates are
        // neither kept nor near; dt < -margin keeps a voxel, |dt| <= margin
        // leaves it to the FP64 decision (which adds a kept one itself) and
        // anything else drops it.
        const int nc = ie - ib + 1, nr = je - jb + 1;
        {
            // offsets by repeated steps from the first column / row
            float dx = voxel_offset(c.dx0, ib, p.finter_x);
#pragma unroll
            for (int k = 0; k < BX; k++) {
                col_terms(dx, ev.ux, cx2[k], cux[k]);
                cx2[k] = k < nc ? cx2[k] : INFINITY;
                dx -= p.finter_x;
            }
            float dy = voxel_offset(c.dy0, jb, p.finter_y);
#pragma unroll
            for (int k = 0; k < BY; k++) {
                row_terms(dy, ev.uy, ev.thr, rt2[k], ruy[k]);
                rt2[k] = k < nr ? rt2[k] : INFINITY;
                dy -= p.finter_y;
            }
        }
        float dt[NC];
        float gap = INFINITY; // the smallest |dt| over the box
#pragma unroll
        for (int jj = 0; jj < BY; jj++) {
#pragma unroll
            for (int ii = 0; ii < BX; ii++) {
                const int k = jj * BX + ii;
                dt[k] = tube_dt(cx2[ii], cux[ii], rt2[jj], ruy[jj]);
                gap = fminf(gap, fabsf(dt[k]));
            }
        }
        float cnt = 0.0f;
        if (gap <= kThrMargin) {
            // rare: the reference's FP64 decision, one voxel at a time, for
            // the candidates within the margin of the cut-off; a kept one is
            // added here (the common pass below only adds dt < -margin)
            uint32_t near = 0;
#pragma unroll
            for (int k = 0; k < NC; k++)
                near |= (uint32_t)(fabsf(dt[k]) <= kThrMargin) << k;
            for (uint32_t m = (uint32_t)opaque((int)near); m; m &= m - 1) {
                const int k = __builtin_ctz(m);
                const int i = ib + k % BX, j = jb + k / BX;
                if (inside_exact(p, ev, s, c.ix0 + i, c.iy0 + j)) {
                    if (KIND == 2) {
                        cnt += 1.0f;
                    } else {
                        const float w = voxel_weight(
                            ev, amp, voxel_dt(ev, voxel_offset(c.dx0, i, p.finter_x),
                                              voxel_offset(c.dy0, j, p.finter_y)));
                        D *q = dst + idx(i, j);
                        if constexpr (LDS != 0)
                            atomicAdd(q, (D)w);
                        else
                            unsafeAtomicAdd(q, w);
                    }
                }
            }
        }
        if (KIND == 2) {
#pragma unroll
            for (int k = 0; k < NC; k++)
                cnt += dt[k] < -kThrMargin ? 1.0f : 0.0f;
            return cnt;
        }
        if constexpr (LDS == 2) {
#pragma unroll
            for (int ii = 0; ii < BX; ii++)
                col[ii] = opaque(idx(ib + ii, jb));
        } else {
#pragma unroll
            for (int jj = 0; jj < BY; jj++)
                row[jj] = LDS ? opaque(idx(ib, jb + jj)) : 0;
        }
#pragma unroll
        for (int k = 0; k < NC; k++) {
            if (dt[k] < -kThrMargin) {
                const float w = voxel_weight(ev, amp, dt[k]);
                D *q = dst + (LDS == 2   ? col[k % BX] + k / BX
                              : LDS == 1 ? row[k / BX] + k % BX
                                         : idx(ib + k % BX, jb + k / BX));
                if constexpr (LDS != 0)
                    atomicAdd(q, (D)w);
                else
                    unsafeAtomicAdd(q, w);
            }
        }
        return 0.0f;
    }
}

// The gather of slice_box<0> from a contiguous op-frame image (unit column
// stride, rows p.sy apart, p.img_bytes long), read through a buffer
// resource: every candidate is loaded, a row's columns at immediate offsets
// from one row offset (a candidate past the image end reads 0; candidate
// indices are never negative, the box starting inside the grid), and the
// dropped candidates' values are zeroed before the sum, so no decision is
// held as a mask across the loads.  Same decisions, weights and summation
// order (a dropped candidate adds fma(w, 0, acc) = acc) as slice_box<0>.
template <int BX, int BY>
__device__ __forceinline__ float slice_gather_frame(const TorParams &p, const Event &ev,
                                                    const Track &t, const Cross &c, float amp,
                                                    int s, int i_lo, int i_hi, int j_lo,
                                                    int j_hi)
{
    constexpr int NC = BX * BY;
    const float eps = 1.0e-4f;
    const int ib0 = (int)ceilf((c.dx0 - t.hx) * p.rfinter_x - eps);
    const int jb0 = (int)ceilf((c.dy0 - t.dymax) * p.rfinter_y - eps);
    const int ib = max(i_lo, ib0), ie = min(i_hi, ib0 + t.nx - 1);
    const int jb = max(j_lo, jb0), je = min(j_hi, jb0 + t.ny - 1);
    float cx2[BX], cux[BX], rt2[BY], ruy[BY];
#pragma unroll
    for (int k = 0; k < BX; k++)
        col_terms(voxel_offset(c.dx0, ib + k, p.finter_x), ev.ux, cx2[k], cux[k]);
#pragma unroll
    for (int k = 0; k < BY; k++)
        row_terms(voxel_offset(c.dy0, jb + k, p.finter_y), ev.uy, ev.thr, rt2[k], ruy[k]);
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
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)p.img, (short)0, p.img_bytes, 0x00020000);
    // 24-bit products (full rate): slice and row indices are non-negative and
    // a frame slice holds fewer than 2^24 voxels
    const unsigned o0 =
        4u * (__umul24((unsigned)s, (unsigned)p.sz) +
              __umul24((unsigned)(c.iy0 + jb), (unsigned)p.sy) + (unsigned)(c.ix0 + ib));
    float v[NC];
#pragma unroll
    for (int jj = 0; jj < BY; jj++) {
        const unsigned orow = o0 + 4u * (unsigned)(jj * p.sy);
#pragma unroll
        for (int ii = 0; ii < BX; ii++)
            v[jj * BX + ii] =
                __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, orow + 4u * ii, 0, 0));
    }
    auto ok = [&](int k) { return (ib + k % BX <= ie) & (jb + k / BX <= je); };
    // a dropped candidate's weight is finite (dt >= -thr), so zeroing its
    // value alone makes its term fma(w, 0, acc) = acc
    auto sum = [&](auto keep) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < NC; k++)
            acc = fmaf(voxel_weight(ev, amp, dt[k]), keep(k) ? v[k] : 0.0f, acc);
        return acc;
    };
    if (gap > kThrMargin)
        return sum([&](int k) { return ok(k) & (dt[k] < -kThrMargin); });
    // rare: the reference's FP64 decision, one voxel at a time
    uint32_t in = 0, near = 0;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        in |= (uint32_t)(ok(k) & (dt[k] < -kThrMargin)) << k;
        near |= (uint32_t)(ok(k) & (fabsf(dt[k]) <= kThrMargin)) << k;
    }
    for (uint32_t m = (uint32_t)opaque((int)near); m; m &= m - 1) {
        const int k = __builtin_ctz(m);
        if (inside_exact(p, ev, s, c.ix0 + ib + k % BX, c.iy0 + jb + k / BX))
            in |= 1u << k;
    }
    return sum([&](int k) { return ((in >> k) & 1u) != 0; });
}

// The gather of a box whose patch lies inside the image (most slices of most
// events): no image clipping, the box kept inside the patch, and the slice's
// amplitude multiplying the slice's sum once instead of every candidate's
// weight.  Returns false, having decided nothing, when the patch is not
// inside, the box starts before the patch, or a candidate lies within
// kThrMargin of the cut-off; the caller then runs the clipped gather, which
// decides those candidates exactly.  The voxels are those of the clipped
// gather: with the patch inside, it drops the candidates past the patch
// (here the box ends on the patch's edge) and those past the ellipse's span
// (ib0 + nx and beyond, jb0 + ny and beyond), which lie outside the padded
// ellipse (dt > kThrMargin) and fail the dt test here.
template <int BX, int BY>
__device__ __forceinline__ bool slice_gather_inner(const TorParams &p, const Event &ev,
                                                   const Track &t, const Cross &c, float amp,
                                                   int s, float &out)
{
    constexpr int NC = BX * BY;
    const float eps = 1.0e-4f;
    const int ib0 = (int)ceilf((c.dx0 - t.hx) * p.rfinter_x - eps);
    const int jb0 = (int)ceilf((c.dy0 - t.dymax) * p.rfinter_y - eps);
    const int P = p.patch;
    const bool inner = (ib0 >= 0) & (jb0 >= 0) & (c.ix0 >= 0) & (c.ix0 <= p.gx - P) &
                       (c.iy0 >= 0) & (c.iy0 <= p.gy - P) & (P >= BX) & (P >= BY);
    if (!inner)
        return false;
    // A box running past the patch's far edge moves back to end on it: a
    // superset of the patch columns (rows) it covered, all inside the patch,
    // so no candidate needs clipping.  The added ones lie before the
    // ellipse's span and fail the dt test (or, within kThrMargin of the
    // cut-off, send the slice to the clipped gather); the kept voxels and
    // their order in the sum are those of the unshifted box.
    const int ib = min(ib0, P - BX), jb = min(jb0, P - BY);
    float cx2[BX], cux[BX], rt2[BY], ruy[BY];
    {
        float dx = fmaf(-(float)ib, p.finter_x, c.dx0);
#pragma unroll
        for (int k = 0; k < BX; k++) {
            col_terms(dx, ev.ux, cx2[k], cux[k]);
            dx -= p.finter_x;
        }
        float dy = fmaf(-(float)jb, p.finter_y, c.dy0);
#pragma unroll
        for (int k = 0; k < BY; k++) {
            row_terms(dy, ev.uy, ev.thr, rt2[k], ruy[k]);
            dy -= p.finter_y;
        }
    }
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)p.img, (short)0, p.img_bytes, 0x00020000);
    // 24-bit products (full rate): slice and row indices are non-negative and
    // a frame slice holds fewer than 2^24 voxels
    const unsigned o0 =
        4u * (__umul24((unsigned)s, (unsigned)p.sz) +
              __umul24((unsigned)(c.iy0 + jb), (unsigned)p.sy) + (unsigned)(c.ix0 + ib));
    float v[NC];
#pragma unroll
    for (int jj = 0; jj < BY; jj++) {
        const unsigned orow = o0 + 4u * (unsigned)(jj * p.sy);
#pragma unroll
        for (int ii = 0; ii < BX; ii++)
            v[jj * BX + ii] =
                __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, orow + 4u * ii, 0, 0));
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
    if (!(gap > kThrMargin))
        return false;
    const float c0 = ev.thr * ev.cw;
    float a = 0.0f;
#pragma unroll
    for (int k = 0; k < NC; k++)
        a = fmaf(__builtin_amdgcn_exp2f(fmaf(dt[k], ev.cw, c0)),
                 dt[k] < -kThrMargin ? v[k] : 0.0f, a);
    out = amp * a;
    return true;
}

// Box shape of an event for a compiled box size NB: 1 = NB x (NB - 1)
// (enough for every event whose ellipse spans at most NB - 1 rows, e.g. all
// C2 events), 2 = NB x NB, 0 = larger (chord walk).
template <int NB>
__device__ __forceinline__ int box_shape(const Track &t)
{
    if (NB <= 0 || t.nx > NB || t.ny > NB)
        return 0;
    return t.ny <= NB - 1 ? 1 : 2;
}

// The box walk of one slice over the strided image (patch voxels inside the
// grid).
template <int KIND, int NB, bool FRAME = false>
__device__ __forceinline__ float slice_box_image(const TorParams &p, const Event &ev,
                                                 const Track &t, const Cross &c, float amp, int s,
                                                 float scale, int shape)
{
    if constexpr (KIND == 0 && FRAME) {
        float r;
        if (shape == 1 ? slice_gather_inner<NB, NB - 1>(p, ev, t, c, amp, s, r)
                       : slice_gather_inner<NB, NB>(p, ev, t, c, amp, s, r))
            return r;
    }
    const int P = p.patch;
    if (c.ix0 <= -P || c.ix0 >= p.gx || c.iy0 <= -P || c.iy0 >= p.gy)
        return 0.0f;
    const int i_lo = c.ix0 < 0 ? -c.ix0 : 0;
    const int i_hi = min(P, p.gx - c.ix0) - 1;
    const int j_lo = c.iy0 < 0 ? -c.iy0 : 0;
    const int j_hi = min(P, p.gy - c.iy0) - 1;
    if constexpr (KIND == 0 && FRAME) {
        if (shape == 1)
            return slice_gather_frame<NB, NB - 1>(p, ev, t, c, amp, s, i_lo, i_hi, j_lo, j_hi);
        return slice_gather_frame<NB, NB>(p, ev, t, c, amp, s, i_lo, i_hi, j_lo, j_hi);
    }
    PatchIdx idx;
    if (KIND == 0) {
        idx.base = s * p.sz + c.ix0 * p.sx + c.iy0 * p.sy;
        idx.si = p.sx;
        idx.sj = p.sy;
    } else {
        idx.base = s * p.oz + c.ix0 * p.ox + c.iy0 * p.oy;
        idx.si = p.ox;
        idx.sj = p.oy;
    }
    if (shape == 1)
        return slice_box<KIND, NB, NB - 1, 0>(p, ev, t, c, amp, s, scale, i_lo, i_hi, j_lo,
                                                  j_hi, idx, p.img, p.out);
    return slice_box<KIND, NB, NB, 0>(p, ev, t, c, amp, s, scale, i_lo, i_hi, j_lo, j_hi,
                                          idx, p.img, p.out);
}

// Walk the crossed slices s_first, s_first + rstep, ... <= s_last of an event
// (the caller clips to [s_lo, s_hi]).  NB > 0: slices go through the NB x NB
// candidate box when the event's box fits, else through the chord walk.
template <int KIND, int NB, bool FRAME = false>
__device__ __forceinline__ float walk_event(const TorParams &p, const Event &ev, const Track &t,
                                            bool tof, float scale, int s_first, int s_last,
                                            int rstep, int s_min = -1)
{
    constexpr int NBB = NB > 1 ? NB : 2;
    const int shape = p.full_patch ? 0 : box_shape<NB>(t);
    const float ts = p.tofk * scale;
    float acc = 0.0f;
    for (int s = s_first; s <= s_last; s += rstep) {
        if (s < s_min)
            continue; // a slice before the event's first (aligned walks)
        Cross c;
        cross_at(p, ev, t, s, c);
        // scatters: the scale rides in amp (one product per slice, not per
        // voxel), with the TOF constant folded into it once; gathers: the TOF
        // constant multiplies the walk's sum once, after the loop
        float amp = KIND == 1 ? (tof ? tof_exp(p, ev, t, s) * ts : scale)
                              : (tof ? tof_exp(p, ev, t, s) : 1.0f);
        if (NB > 0 && shape)
            acc += slice_box_image<KIND, NBB, FRAME>(p, ev, t, c, amp, s, scale, shape);
        else if (p.full_patch)
            acc += patch_walk<KIND, true>(p, ev, t, c, amp, s, scale);
        else
            acc += patch_walk<KIND, false>(p, ev, t, c, amp, s, scale);
    }
    return KIND == 0 && tof ? acc * p.tofk : acc;
}

// Sum over the LPE consecutive lanes that share an event.  Every lane of the
// wave must reach this call.
template <int LPE>
__device__ __forceinline__ float group_sum(float v)
{
#pragma unroll
    for (int m = 1; m < LPE; m <<= 1)
        v += __shfl_xor(v, m, 64);
    return v;
}

// One event of mode MODE (the body of the reference's per-slice kernels,
// tor.cu.cc:244-346, for all of the event's slices), shared by LPE lanes:
// lane `sub` takes every LPE-th crossed slice and partial sums are combined
// across the lanes.  Every lane of the wave calls this; lanes with
// valid == false contribute nothing.  Returns the event's non-zero count on
// its first lane in kCount mode.
template <int MODE, int NB, int LPE>
__device__ __forceinline__ int run_event(const TorParams &p, int64_t e, int sub, bool valid)
{
    Event ev;
    Track t;
    bool any = valid ? load_event<MODE>(p, e, ev, t) : false;
    const bool tof = (MODE != kMap && MODE != kCount) && p.use_tof;
    if (MODE == kProj) {
        float part = any ? walk_event<0, NB>(p, ev, t, tof, 0.0f, ev.s_lo + sub, ev.s_hi, LPE) : 0.0f;
        float pv = group_sum<LPE>(part);
        if (valid && sub == 0)
            p.out[e * p.out_step] = pv;
    } else if (MODE == kBack) {
        float v = valid ? p.val[e] : 0.0f;
        if (any && v > 1e-7)
            walk_event<1, NB>(p, ev, t, tof, 1.0f / v, ev.s_lo + sub, ev.s_hi, LPE);
    } else if (MODE == kMap) {
        if (any)
            walk_event<1, NB>(p, ev, t, false, 1.0f / p.val[e], ev.s_lo + sub, ev.s_hi, LPE);
    } else if (MODE == kAdj) {
        float g = valid ? p.val[e] : 0.0f;
        if (any && g != 0.0f)
            walk_event<1, NB>(p, ev, t, tof, g, ev.s_lo + sub, ev.s_hi, LPE);
    } else if (MODE == kOsem) {
        float part = any ? walk_event<0, NB>(p, ev, t, tof, 0.0f, ev.s_lo + sub, ev.s_hi, LPE) : 0.0f;
        float pv = group_sum<LPE>(part);
        if (any && pv > 1e-7)
            walk_event<1, NB>(p, ev, t, tof, 1.0f / pv, ev.s_lo + sub, ev.s_hi, LPE);
    } else { // kCount
        float part = any ? walk_event<2, NB>(p, ev, t, false, 0.0f, ev.s_lo + sub, ev.s_hi, LPE) : 0.0f;
        int cnt = (int)group_sum<LPE>(part);
        if (valid && sub == 0) {
            if (p.nnz)
                p.nnz[e] = cnt;
            return cnt;
        }
    }
    return 0;
}

template <int MODE, int NB, int LPE>
__global__ __launch_bounds__(256) void tor_kernel(TorParams p)
{
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t e = gid / LPE;
    int sub = (int)(gid % LPE);
    int cnt = run_event<MODE, NB, LPE>(p, e, sub, e < p.n);
    if (MODE == kCount && p.nnz_total) {
        // Wave reduction, then one 64-bit atomic per wave.
        unsigned long long c = (unsigned long long)cnt;
        for (int off = 32; off > 0; off >>= 1)
            c += __shfl_down(c, off, 64);
        if ((threadIdx.x & 63) == 0 && c)
            atomicAdd(p.nnz_total, c);
    }
}

// Slab schedule.  The slices are cut into nslab contiguous slabs and every
// event is visited once per slab, by a lane that walks only the event's
// slices inside that slab.  With interleave, workgroup b works on slab
// b % nslab: the dispatcher hands consecutive workgroups to consecutive XCDs,
// so with nslab a multiple of 8 each XCD gathers from and scatters into its
// own slabs only, and a 16 MB image splits into L2-sized pieces instead of
// every XCD's L2 cycling through all of it.  Without interleave each slab gets
// a contiguous range of workgroups (same work, no XCD affinity).
//   kProj: part[slab * n + e] = the event's projection over the slab;
//   kOsem: the scatter half of the OSEM step, with the projection
//          sum_k part[k * n + e] (slab order, so deterministic);
//   kBack / kMap / kAdj: the scatter over the slab.
__device__ __forceinline__ void slab_of(const TorParams &p, int nslab, int interleave, int &slab,
                                        int64_t &chunk)
{
    if (interleave) {
        slab = (int)(blockIdx.x % (unsigned)nslab);
        chunk = blockIdx.x / (unsigned)nslab;
    } else {
        unsigned nchunk = gridDim.x / (unsigned)nslab;
        slab = (int)(blockIdx.x / nchunk);
        chunk = blockIdx.x % nchunk;
    }
}

template <int MODE, int NB>
__global__ __launch_bounds__(256) void tor_slab_kernel(TorParams p, float *part, int nslab,
                                                       int interleave)
{
    int slab;
    int64_t chunk;
    slab_of(p, nslab, interleave, slab, chunk);
    const int64_t e = chunk * blockDim.x + threadIdx.x;
    const bool valid = e < p.n;
    const int sa = (int)((int64_t)slab * p.gz / nslab);
    const int sb = (int)((int64_t)(slab + 1) * p.gz / nslab) - 1;
    Event ev;
    Track t;
    bool any = valid ? load_event<MODE>(p, e, ev, t) : false;
    int s0 = 0, s1 = -1;
    if (any) {
        s0 = max(ev.s_lo, sa);
        s1 = min(ev.s_hi, sb);
        any = s0 <= s1;
    }
    const bool tof = MODE != kMap && p.use_tof;
    if (MODE == kProj) {
        float v = any ? walk_event<0, NB>(p, ev, t, tof, 0.0f, s0, s1, 1) : 0.0f;
        if (valid)
            part[(int64_t)slab * p.n + e] = v;
    } else {
        float scale = 0.0f;
        bool go = any;
        if (MODE == kOsem) {
            float v = 0.0f;
            if (valid)
                for (int k = 0; k < nslab; k++)
                    v += part[(int64_t)k * p.n + e];
            go = go && v > 1e-7;
            scale = 1.0f / v;
        } else if (MODE == kBack) {
            float v = valid ? p.val[e] : 0.0f;
            go = go && v > 1e-7;
            scale = 1.0f / v;
        } else if (MODE == kMap) {
            scale = valid ? 1.0f / p.val[e] : 0.0f;
        } else { // kAdj
            float g = valid ? p.val[e] : 0.0f;
            go = go && g != 0.0f;
            scale = g;
        }
        if (go)
            walk_event<1, NB>(p, ev, t, tof, scale, s0, s1, 1);
    }
}

// out[e] = sum over slabs of part[k * n + e], slab order.
__global__ __launch_bounds__(256) void tor_slab_sum(const float *part, int nslab, int64_t n,
                                                    float *out)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n)
        return;
    float v = 0.0f;
    for (int k = 0; k < nslab; k++)
        v += part[(int64_t)k * n + e];
    out[e] = v;
}

// Host part of the reference setup (tor.cu.cc:364-375), in double.
int setup(TorParams &p, const float *lors, int64_t n, int64_t ld, const srfx_grid &g,
          float kernel_width, int geo, float tof_bin, float tof_sigma2)
{
    if (n < 0 || (n > 0 && (!lors || ld < n)) || !grid_ok(g))
        return SRFX_EINVAL;
    if (!(kernel_width > 0.0f) || !(kernel_width < 1.0e6f))
        return SRFX_EINVAL;
    p = TorParams{};
    p.out_step = 1;
    p.lors = lors;
    p.n = n;
    p.ld = ld;
    p.gx = g.shape[0];
    p.gy = g.shape[1];
    p.gz = g.shape[2];
    double lx = g.size[0], ly = g.size[1], lz = g.size[2];
    p.inter_x = lx / (double)p.gx;
    p.inter_y = ly / (double)p.gy;
    p.inter_z = lz / (double)p.gz;
    p.rinter_x = 1.0 / p.inter_x;
    p.rinter_y = 1.0 / p.inter_y;
    p.rinter_z = 1.0 / p.inter_z;
    p.finter_x = (float)p.inter_x;
    p.finter_y = (float)p.inter_y;
    p.rfinter_x = (float)(1.0 / p.inter_x);
    p.rfinter_y = (float)(1.0 / p.inter_y);
    p.l_bound = (double)g.center[0] - lx / 2;
    p.b_bound = (double)g.center[1] - ly / 2;
    p.zbase = (double)g.center[2] - (lz - p.inter_z) / 2;
    double kw = kernel_width;
    p.tor_sigma2 = kw * kw / 36;
    double ps = std::ceil((std::sqrt(2.0) * kw + p.inter_z) / p.inter_x);
    if (!(ps >= 1.0) || ps > kMaxPatch)
        return SRFX_EUNSUPPORTED;
    p.patch = (int)ps;
    p.half = (int)((unsigned)p.patch / 2);
    p.dx_half = ((double)p.half - 0.5) * p.inter_x;
    p.dy_half = ((double)p.half - 0.5) * p.inter_y;
    p.fdx_half = (float)p.dx_half;
    p.fdy_half = (float)p.dy_half;
    p.geo = geo ? 1 : 0;
    p.use_tof = tof_sigma2 < kTofThreshold ? 1 : 0;
    if (p.use_tof) {
        double tse = (double)tof_sigma2 + (double)tof_bin * (double)tof_bin / 12;
        if (!(tse > 0.0) || !(tof_sigma2 > 0.0f))
            return SRFX_EINVAL;
        p.tse_inv = (float)(1.0 / tse);
        p.tofk = (float)((double)tof_bin / std::sqrt(2.0 * M_PI * (double)tof_sigma2));
    }
    // Candidate box: the compiled size that holds the box of every event
    // whose slice axis is its dominant direction (as in the axis groups of
    // on_tor_lors.py): there the ellipse's half extents are at most
    // sqrt(2 tm).  Events with a larger box take the chord walk.
    {
        double tm = 9.0 * p.tor_sigma2 * (1.0 + 1.0e-5) + 2.0 * kThrMargin;
        double h = std::sqrt(2.0 * tm);
        int nx = (int)std::floor(2.0 * h / p.inter_x + 2.0e-4) + 1;
        int ny = (int)std::floor(2.0 * h / p.inter_y + 2.0e-4) + 1;
        int n = nx > ny ? nx : ny;
        p.box = n <= 3 ? 3 : (n <= 4 ? 4 : 0);
    }
    Strides s = grid_strides(g);
    p.sx = p.ox = (int)s.x;
    p.sy = p.oy = (int)s.y;
    p.sz = p.oz = (int)s.z;
    return SRFX_OK;
}

template <int MODE, int NB, int LPE>
int launch_pg(const TorParams &p, int block, hipStream_t st)
{
    hipLaunchKernelGGL((tor_kernel<MODE, NB, LPE>), dim3(blocks_for(p.n * LPE, block)),
                       dim3(block), 0, st, p);
    return launch_status();
}

template <int MODE, int LPE>
int launch_lpe(const TorParams &p, int block, hipStream_t st)
{
    switch (p.full_patch ? 0 : p.box) {
    case 3:
        return launch_pg<MODE, 3, LPE>(p, block, st);
    case 4:
        return launch_pg<MODE, 4, LPE>(p, block, st);
    default:
        return launch_pg<MODE, 0, LPE>(p, block, st);
    }
}

template <int MODE, int NB>
int launch_slab_pg(const TorParams &p, float *part, int nslab, int interleave, int block,
                   hipStream_t st)
{
    unsigned chunks = blocks_for(p.n, block);
    hipLaunchKernelGGL((tor_slab_kernel<MODE, NB>), dim3(chunks * (unsigned)nslab), dim3(block),
                       0, st, p, part, nslab, interleave);
    return launch_status();
}

// Slab schedule (flags field SRFX_SLAB_*): kProj writes the per-slab partial
// projections to part, kOsem / kBack / kMap / kAdj scatter (kOsem reading
// part).  Returns SRFX_EUNSUPPORTED for schedules without slabs.
template <int MODE>
int launch_slab(const TorParams &p, float *part, uint32_t flags, hipStream_t st)
{
    int nslab, interleave;
    if (!slab_schedule(flags, nslab, interleave))
        return SRFX_EUNSUPPORTED;
    static const int kBlock[4] = {256, 64, 128, 256};
    const int block = kBlock[(flags >> SRFX_BLOCK_SHIFT) & 3u];
    if (p.n == 0)
        return SRFX_OK;
    switch (p.full_patch ? 0 : p.box) {
    case 3:
        return launch_slab_pg<MODE, 3>(p, part, nslab, interleave, block, st);
    case 4:
        return launch_slab_pg<MODE, 4>(p, part, nslab, interleave, block, st);
    default:
        return launch_slab_pg<MODE, 0>(p, part, nslab, interleave, block, st);
    }
}

int launch_slab_sum(const float *part, int nslab, int64_t n, float *out, hipStream_t st)
{
    if (n == 0)
        return SRFX_OK;
    hipLaunchKernelGGL(tor_slab_sum, dim3(blocks_for(n, 256)), dim3(256), 0, st, part, nslab, n,
                       out);
    return launch_status();
}

// Launch schedule from the flags: lanes per event (1, 4, 16) and block size
// (64, 128, 256).  Kernels are specialised for the candidate box size
// (3 or 4); the full-patch verification path and larger boxes use the
// generic loops.
template <int MODE>
int launch(const TorParams &p, uint32_t flags, hipStream_t st)
{
    if (p.n == 0)
        return SRFX_OK;
    static const int kBlock[4] = {256, 64, 128, 256};
    const int block = kBlock[(flags >> SRFX_BLOCK_SHIFT) & 3u];
    switch ((flags >> SRFX_LPE_SHIFT) & 3u) {
    case 2:
        return launch_lpe<MODE, 4>(p, block, st);
    case 3:
        return launch_lpe<MODE, 16>(p, block, st);
    default:
        return launch_lpe<MODE, 1>(p, block, st);
    }
}

srfx_tor_model model_no_tof(float kw, int geo)
{
    srfx_tor_model m;
    m.kernel_width = kw;
    m.geo_sigma2_flag = geo;
    m.tof_bin = 0.0f;
    m.tof_sigma2 = 1.0e6f;
    return m;
}

} // namespace
} // namespace srfx
