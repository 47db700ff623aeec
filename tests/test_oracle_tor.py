"""Pins the TOR oracle (FP64 restatement of tor.cu.cc) with checks that do
not reuse its formulas:

* brute force on a small grid: crossing point from the parametric line, the
  squared distance of each patch voxel centre to the 3-D LOR line from the
  cross product |(m - p1) x u|^2, the TOF term from the signed distance along
  the LOR between the crossing point and the TOF centre;
* closed form: a LOR along the slice axis through voxel centres sees only the
  centre voxel of every slice (the neighbours sit exactly on the 3-sigma cut,
  which is strict), with the TOF Gaussian of the slice's axial offset;
* adjoint identity, the 1e-7 backprojection threshold, Maplors == non-TOF
  backprojection, geo_sigma2_flag with a unit factor == flag off;
* the FP32 port's distance to FP64 (the reference's own FP32 error).
"""
import math

import numpy as np
import pytest

from oracle import lib as O
from tests import helpers as H

SMALL = ((12, 10, 8), (0.5, -0.25, 0.3), (24.0, 20.0, 17.6))


def brute_tor(lors, img, grid, center, size, kw, geo, tof_bin, tof_sigma2):
    g = np.array(grid)
    c = np.array(center, np.float32).astype(np.float64)
    s = np.array(size, np.float32).astype(np.float64)
    inter = s / g
    lb = c - s / 2
    kw = float(np.float32(kw))
    sig2 = kw * kw / 36
    P = math.ceil((math.sqrt(2) * kw + inter[2]) / inter[0])
    zs = c[2] - (s[2] - inter[2]) / 2 + np.arange(g[2]) * inter[2]
    tof = tof_sigma2 < 405758.0
    tb, ts2 = float(np.float32(tof_bin)), float(np.float32(tof_sigma2))
    out = np.zeros(lors.shape[1])
    for e in range(lors.shape[1]):
        col = lors[:, e].astype(np.float64)
        p1, p2, pc, f = col[0:3], col[3:6], col[6:9], col[9]
        d = p2 - p1
        u = d / np.linalg.norm(d)
        sc = sig2 * (f if geo else 1.0)
        acc = 0.0
        for si, z in enumerate(zs):
            if d[2] == 0:
                continue
            t = (z - p1[2]) / d[2]
            if not (0 < t < 1):
                continue
            X = p1 + t * d
            ix0 = int((X[0] - lb[0]) / inter[0]) - P // 2
            iy0 = int((X[1] - lb[1]) / inter[1]) - P // 2
            for iy in range(iy0, iy0 + P):
                for ix in range(ix0, ix0 + P):
                    if not (0 <= ix < g[0] and 0 <= iy < g[1]):
                        continue
                    m = np.array([lb[0] + (ix + 0.5) * inter[0], lb[1] + (iy + 0.5) * inter[1], z])
                    dist2 = float(np.sum(np.cross(m - p1, u) ** 2))
                    if dist2 >= 9 * sc:
                        continue
                    w = math.exp(-dist2 / (2 * sc))
                    if tof:
                        along = float(np.dot(X - pc, u))
                        w *= tb * math.exp(-(along * along - dist2) / (2 * (ts2 + tb * tb / 12))) \
                            / math.sqrt(2 * math.pi * ts2)
                    acc += w * float(img[si, iy, ix])
        out[e] = acc
    return out


def small_lors(n, seed):
    """LORs crossing the SMALL box along its slice axis, TOF centre on the
    segment, sigma2 factor in [0.3, 1.7]."""
    rng = np.random.default_rng(seed)
    c = np.array(SMALL[1])
    s = np.array(SMALL[2])
    lo, hi = c - s / 2, c + s / 2
    p1 = np.column_stack([rng.uniform(lo[0] - 2, hi[0] + 2, n), rng.uniform(lo[1] - 2, hi[1] + 2, n),
                          rng.uniform(lo[2] - 8, lo[2] + 3, n)])
    p2 = p1 + np.column_stack([rng.uniform(-8, 8, n), rng.uniform(-8, 8, n),
                               rng.uniform(s[2] - 3, s[2] + 8, n)])
    pc = p1 + rng.uniform(0.2, 0.8, n)[:, None] * (p2 - p1)
    f = rng.uniform(0.3, 1.7, n)
    return np.vstack([p1.T, p2.T, pc.T, f[None]]).astype(np.float32)


@pytest.mark.parametrize("tof,geo", [(True, False), (False, True), (True, True)])
def test_brute_force(tof, geo):
    grid, center, size = SMALL
    lors = small_lors(150, seed=21)
    img = H.image(grid, seed=4)
    ts2 = 60.0 if tof else 1.0e6
    ref = brute_tor(lors, img, grid, center, size, 4.0, geo, 3.0, ts2)
    got = O.tor_project(lors, img, grid, center, size, 4.0, geo, 3.0, ts2, prec="f64")
    assert np.count_nonzero(ref) > 100
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12)


def test_axis_line_closed_form():
    """LOR along the slice axis through the centre of voxel (ix, iy):
    d2 = 0 for that voxel and = pitch^2 = 9 sigma^2 (excluded, strict) for
    its neighbours at kernel_width = 3 * pitch... here pitch 2 mm, kw = 4 mm:
    9 sigma^2 = kw^2/4 = 4 = pitch^2."""
    grid, center, size = SMALL
    c = np.array(center, np.float32).astype(np.float64)
    s = np.array(size, np.float32).astype(np.float64)
    inter = s / np.array(grid)
    ix, iy = 5, 4
    x = c[0] - s[0] / 2 + (ix + 0.5) * inter[0]
    y = c[1] - s[1] / 2 + (iy + 0.5) * inter[1]
    zlo, zhi = c[2] - s[2], c[2] + s[2]
    zc = c[2] + 3.1
    lors = np.array([[x], [y], [zlo], [x], [y], [zhi], [x], [y], [zc], [1.0]], np.float32)
    img = H.image(grid, seed=9)
    # non-TOF: sum of the centre voxel over all slices
    p = O.tor_project(lors, img, grid, center, size, 4.0, False, 3.0, 1.0e6, prec="f64")[0]
    assert p == pytest.approx(float(img[:, iy, ix].astype(np.float64).sum()), rel=1e-12)
    # TOF: Gaussian of the slice offset from the TOF centre
    tb, ts2 = 3.0, 60.0
    zs = c[2] - (s[2] - inter[2]) / 2 + np.arange(grid[2]) * inter[2]
    zc32 = float(np.float32(zc))
    wt = tb * np.exp(-0.5 * (zc32 - zs) ** 2 / (ts2 + tb * tb / 12)) / np.sqrt(2 * np.pi * ts2)
    p = O.tor_project(lors, img, grid, center, size, 4.0, False, tb, ts2, prec="f64")[0]
    assert p == pytest.approx(float(np.sum(wt * img[:, iy, ix])), rel=1e-12)
    # count: one voxel per slice
    cnt = O.tor_count(lors, grid, center, size, 4.0, False, prec="f64")[0]
    assert cnt == grid[2]


def test_adjoint_threshold_map_geo():
    lors = H.tor_lors(800, seed=22)
    g, c, s = H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 1, g[::-1]).astype(np.float32)
    y = rng.normal(0, 1, lors.shape[1]).astype(np.float32)
    args = (g, c, s, H.KW, False, H.TOF_BIN, H.TOF_SIGMA2)
    Ax = O.tor_project(lors, x, *args, prec="f64")
    ATy = O.tor_backproject_adjoint(lors, y, *args, prec="f64")
    lhs, rhs = Ax @ y.astype(np.float64), np.sum(x.astype(np.float64) * ATy)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # threshold: v <= 1e-7 contributes nothing
    v = rng.uniform(0.5, 2.0, lors.shape[1]).astype(np.float32)
    v[::2] = 1e-8
    b = O.tor_backproject(lors, v, *args, prec="f64")
    gi = np.where(v > 1e-7, 1.0 / v.astype(np.float64), 0.0).astype(np.float32)
    a = O.tor_backproject_adjoint(lors, gi, *args, prec="f64")
    assert H.rel_l2(b, a) < 1e-7
    # Maplors == non-TOF backprojection (same weights in FP64)
    l7 = np.ascontiguousarray(np.vstack([lors[0:6], lors[9:10]]))
    ones = np.ones(lors.shape[1], np.float32)
    m = O.tor_map(l7, ones, g, c, s, H.KW, True, prec="f64")
    b0 = O.tor_backproject(lors, ones, g, c, s, H.KW, True, 0.0, 1.0e6, prec="f64")
    assert H.rel_l2(m, b0) < 1e-13
    # geo flag with factor 1 == flag off
    l1 = lors.copy()
    l1[9] = 1.0
    p_on = O.tor_project(l1, x, g, c, s, H.KW, True, H.TOF_BIN, H.TOF_SIGMA2, prec="f64")
    p_off = O.tor_project(l1, x, g, c, s, H.KW, False, H.TOF_BIN, H.TOF_SIGMA2, prec="f64")
    np.testing.assert_array_equal(p_on, p_off)


def test_fp32_port_error_level():
    """The reference's FP32 arithmetic deviates from FP64 by ~1e-5 on this
    narrow kernel (crossing-point rounding); the port reproduces that level."""
    lors = H.tor_lors(3000, seed=23)
    g, c, s = H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE
    x = H.image(g, seed=3)
    args = (g, c, s, H.KW, False, H.TOF_BIN, H.TOF_SIGMA2)
    a = O.tor_project(lors, x, *args, prec="f64")
    b = O.tor_project(lors, x, *args, prec="f32")
    err = H.rel_l2(b, a)
    assert 1e-7 < err < 1e-4


def brute_tor_count(lors, grid, center, size, kw, geo):
    """Per-event number of kept (event, voxel) pairs from the brute-force
    geometry above (parametric crossing, cross-product distance): strictly
    crossed slice, voxel in the grid and in the P x P patch, d^2 < 9 sigma_c^2."""
    g = np.array(grid)
    c = np.array(center, np.float32).astype(np.float64)
    s = np.array(size, np.float32).astype(np.float64)
    inter = s / g
    lb = c - s / 2
    kw = float(np.float32(kw))
    sig2 = kw * kw / 36
    P = math.ceil((math.sqrt(2) * kw + inter[2]) / inter[0])
    zs = c[2] - (s[2] - inter[2]) / 2 + np.arange(g[2]) * inter[2]
    out = np.zeros(lors.shape[1], np.int64)
    for e in range(lors.shape[1]):
        col = lors[:, e].astype(np.float64)
        p1, p2, f = col[0:3], col[3:6], col[9]
        d = p2 - p1
        u = d / np.linalg.norm(d)
        sc = sig2 * (f if geo else 1.0)
        if d[2] == 0:
            continue
        for z in zs:
            t = (z - p1[2]) / d[2]
            if not (0 < t < 1):
                continue
            X = p1 + t * d
            ix0 = int((X[0] - lb[0]) / inter[0]) - P // 2
            iy0 = int((X[1] - lb[1]) / inter[1]) - P // 2
            for iy in range(max(iy0, 0), min(iy0 + P, g[1])):
                for ix in range(max(ix0, 0), min(ix0 + P, g[0])):
                    m = np.array([lb[0] + (ix + 0.5) * inter[0], lb[1] + (iy + 0.5) * inter[1], z])
                    if float(np.sum(np.cross(m - p1, u) ** 2)) < 9 * sc:
                        out[e] += 1
    return out


# the SMALL box at 2.0 x 2.0 x 2.2 mm (patch 4) and at 1.6 x 1.6 x 1.7 mm (patch 5)
SMALL_16 = ((12, 10, 8), (0.5, -0.25, 0.3), (19.2, 16.0, 13.6))


@pytest.mark.parametrize("box", [SMALL, SMALL_16])
@pytest.mark.parametrize("geo", [False, True])
def test_count_brute_force_support(box, geo):
    """tor_count_f64 equals the brute-force support, event by event, and
    every counted pair carries a non-zero weight in the projection of a
    ones image (the support of the projection)."""
    grid, center, size = box
    lors = small_lors(120, seed=31)
    if box is SMALL_16:
        # the same LOR shapes scaled into the smaller box
        c = np.array(center, np.float32)[:, None]
        k = np.array([0.8, 0.8, 0.7727], np.float32)[:, None]
        for r in (0, 3, 6):
            lors[r:r + 3] = (lors[r:r + 3] - c) * k + c
    ref = brute_tor_count(lors, grid, center, size, 4.0, geo)
    got = O.tor_count(lors, grid, center, size, 4.0, geo, prec="f64")
    assert ref.sum() > 500
    np.testing.assert_array_equal(got, ref)
    # an event with a zero count projects a ones image to zero, and back
    ones = np.ones(grid[::-1], np.float32)
    p = O.tor_project(lors, ones, grid, center, size, 4.0, geo, 3.0, 1.0e6, prec="f64")
    assert np.all((p > 0) == (ref > 0))
