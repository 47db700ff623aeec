"""Pins the Siddon oracle (FP64 restatement of siddon.cu.cc) with checks that
do not reuse its formulas:

* chord length: for a ones image without TOF, p * L^2 equals the length of the
  ray inside the box, computed here by the slab method;
* dense brute force on tiny grids: every system-matrix entry a_ij * L^2 equals
  the length of the ray inside voxel j, from a per-voxel slab intersection;
* TOF: each entry equals l/L^2 * bin * exp(-t^2 / 2s2) / sqrt(2 pi s2) with
  t = L/2 - tof - (segment midpoint distance from p1) and
  s2 = tof_sigma2 + (l^2 + bin^2)/12, kept when t^2/s2 < 9, using the
  intersection intervals from the slab method;
* adjoint <Ax, y> = <x, A^T y>, and the backprojection / map thresholds.
"""
import numpy as np
import pytest

from oracle import lib as O
from tests import helpers as H


def slab(p1, p2, lo, hi):
    """(t_in, t_out) of the segment p1 -> p2 (t in mm from p1) inside the
    axis-aligned box [lo, hi], or None."""
    d = p2 - p1
    L = np.linalg.norm(d)
    u = d / L
    t0, t1 = 0.0, L
    for k in range(3):
        if u[k] == 0.0:
            if p1[k] < lo[k] or p1[k] > hi[k]:
                return None
            continue
        a = (lo[k] - p1[k]) / u[k]
        b = (hi[k] - p1[k]) / u[k]
        t0 = max(t0, min(a, b))
        t1 = min(t1, max(a, b))
    return (t0, t1) if t0 < t1 else None


def dense(lors, grid, center, size, tof_bin=None, tof_sigma2=None):
    """Dense A[i, z, y, x] from per-voxel slab intersections."""
    grid = np.array(grid)
    c = np.array(center, np.float32).astype(np.float64)  # the ops take float32 geometry
    s = np.array(size, np.float32).astype(np.float64)
    v = s / grid
    lo0 = c - s / 2
    n = lors.shape[1]
    A = np.zeros((n, grid[2], grid[1], grid[0]))
    for i in range(n):
        p1 = lors[0:3, i].astype(np.float64)
        p2 = lors[3:6, i].astype(np.float64)
        L = np.linalg.norm(p2 - p1)
        for iz in range(grid[2]):
            for iy in range(grid[1]):
                for ix in range(grid[0]):
                    lo = lo0 + v * np.array([ix, iy, iz])
                    r = slab(p1, p2, lo, lo + v)
                    if r is None:
                        continue
                    l = r[1] - r[0]
                    w = l / L ** 2
                    if tof_sigma2 is not None:
                        t = L / 2 - float(lors[6, i]) - 0.5 * (r[0] + r[1])
                        s2 = tof_sigma2 + (l * l + tof_bin * tof_bin) / 12
                        if t * t / s2 >= 9:
                            continue
                        w *= tof_bin * np.exp(-0.5 * t * t / s2) / np.sqrt(2 * np.pi * s2)
                    A[i, iz, iy, ix] = w
    return A


TINY = ((5, 4, 3), (0.3, -0.2, 0.1), (10.0, 8.0, 6.6))


def tiny_rays(n, seed, faces=False):
    rng = np.random.default_rng(seed)
    grid, center, size = TINY
    c = np.array(center)
    out = []
    while len(out) < n:
        a = c + rng.normal(size=3) * 12
        b = c + rng.normal(size=3) * 12
        if np.linalg.norm(b - a) > 20:
            out.append(np.concatenate([a, b, [rng.normal(0, 4)]]))
    # axis-aligned rays through voxel centres and along voxel faces
    v = np.array(size) / np.array(grid)
    lo = c - np.array(size) / 2
    for k in range(3):
        p = lo + v * np.array([1.5, 2.5, 1.5])
        q = p.copy()
        p[k] -= 20
        q[k] += 20
        out.append(np.concatenate([p, q, [0.0]]))
    if faces:  # rays along voxel faces: the slab brute force is ambiguous there
        out = []
        for k in range(3):
            p2 = lo + v * np.array([1.5, 1.5, 1.5])
            p2[(k + 1) % 3] = lo[(k + 1) % 3] + 2 * v[(k + 1) % 3]
            q2 = p2.copy()
            p2[k] -= 20
            q2[k] += 20
            out.append(np.concatenate([p2, q2, [0.0]]))
    return np.array(out).T.astype(np.float32)


def test_chord_length_ones_image():
    lors = H.siddon_lors(400, seed=11, special=True)
    grid, center, size = H.SID_GRID, H.SID_CENTER, H.SID_SIZE
    ones = np.ones(grid[::-1], np.float32)
    p = O.siddon_project(lors, ones, grid, center, size, 0.0, 1.0e6, prec="f64")
    c = np.array(center, np.float32).astype(np.float64)
    s = np.array(size, np.float32).astype(np.float64)
    hit = 0
    for i in range(lors.shape[1]):
        p1 = lors[0:3, i].astype(np.float64)
        p2 = lors[3:6, i].astype(np.float64)
        L = np.linalg.norm(p2 - p1)
        r = slab(p1, p2, c - s / 2, c + s / 2)
        chord = 0.0 if r is None else r[1] - r[0]
        hit += chord > 0
        assert p[i] * L * L == pytest.approx(chord, rel=1e-9, abs=1e-9), i
    assert hit > 100


@pytest.mark.parametrize("tof", [False, True])
def test_dense_brute_force(tof):
    grid, center, size = TINY
    lors = tiny_rays(30, seed=5)
    bin_, s2 = 2.0, 9.0
    A = dense(lors, grid, center, size, bin_ if tof else None, s2 if tof else None)
    ts2 = s2 if tof else 1.0e6
    n = lors.shape[1]
    nvox = int(np.prod(grid))
    rng = np.random.default_rng(3)
    x = rng.uniform(0.5, 2.0, grid[::-1]).astype(np.float32)
    p = O.siddon_project(lors, x, grid, center, size, bin_, ts2, prec="f64")
    ref = A.reshape(n, nvox) @ x.reshape(-1).astype(np.float64)
    np.testing.assert_allclose(p, ref, rtol=1e-9, atol=1e-12)
    # every column of A^T through the adjoint with one-hot g
    for i in range(0, n, 3):
        g = np.zeros(n, np.float32)
        g[i] = 1.0
        col = O.siddon_backproject_adjoint(lors, g, grid, center, size, bin_, ts2, prec="f64")
        np.testing.assert_allclose(col, A[i], rtol=1e-9, atol=1e-12)


def test_face_ray_goes_to_upper_voxel():
    """A ray lying on a voxel face is attributed to the voxel whose index is
    (int)(pt / v_res) on that axis (siddon.cu.cc:237), i.e. the upper one."""
    grid, center, size = TINY
    lors = tiny_rays(0, seed=0, faces=True)
    c = np.array(center, np.float32).astype(np.float64)
    sz = np.array(size, np.float32).astype(np.float64)
    v = sz / np.array(grid)
    for i in range(lors.shape[1]):
        k = i  # ray along axis k, on a face of axis (k + 1) % 3
        g = np.zeros(lors.shape[1], np.float32)
        g[i] = 1.0
        col = O.siddon_backproject_adjoint(lors, g, grid, center, size, 0.0, 1.0e6, prec="f64")
        L = np.linalg.norm(lors[3:6, i].astype(np.float64) - lors[0:3, i])
        nz = np.argwhere(col > 0)  # (z, y, x)
        assert len(nz) == grid[k]
        np.testing.assert_allclose(col[col > 0] * L * L, v[k], rtol=1e-9)
        face_axis = (k + 1) % 3
        idx = {0: 2, 1: 1, 2: 0}[face_axis]  # column of nz for that axis
        assert np.all(nz[:, idx] == 2)


def test_nnz_matches_dense_support():
    grid, center, size = TINY
    lors = tiny_rays(30, seed=6)
    A = dense(lors, grid, center, size, 2.0, 9.0)
    cnt = O.siddon_count(lors, grid, center, size, 2.0, 9.0, prec="f64")
    np.testing.assert_array_equal(cnt, (A.reshape(A.shape[0], -1) > 0).sum(1))


def test_adjoint_and_thresholds():
    lors = H.siddon_lors(500, seed=12)
    g, c, s = H.SID_GRID, H.SID_CENTER, H.SID_SIZE
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 1, g[::-1]).astype(np.float32)
    y = rng.normal(0, 1, lors.shape[1]).astype(np.float32)
    Ax = O.siddon_project(lors, x, g, c, s, 6.0, 400.0, prec="f64")
    ATy = O.siddon_backproject_adjoint(lors, y, g, c, s, 6.0, 400.0, prec="f64")
    lhs, rhs = Ax @ y.astype(np.float64), np.sum(x.astype(np.float64) * ATy)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # ratio backprojection == adjoint of 1/v on v > 0, nothing for v <= 0
    v = rng.uniform(0.5, 2.0, lors.shape[1]).astype(np.float32)
    v[::3] = 0.0
    v[1::7] = -1.0
    b = O.siddon_backproject(lors, v, g, c, s, 6.0, 400.0, prec="f64")
    gi = np.where(v > 0, 1.0 / v.astype(np.float64), 0.0).astype(np.float32)
    a = O.siddon_backproject_adjoint(lors, gi, g, c, s, 6.0, 400.0, prec="f64")
    assert H.rel_l2(b, a) < 1e-7
    # map == non-TOF ratio backprojection (weight 1/(L*L) vs 1/L/L in double)
    m = O.siddon_map(lors, v, g, c, s, prec="f64")
    b0 = O.siddon_backproject(lors, v, g, c, s, 0.0, 1.0e6, prec="f64")
    assert H.rel_l2(m, b0) < 1e-12


def test_tof_threshold_default_keeps_tof_on():
    """TOF(1e4, 1e4) gives tof_sigma2 = 405757.98, which as a float is below
    405758: the reference keeps the TOF branch on (srfapp_new.py:244-251)."""
    from oracle.preprocess import tof_info
    tb, ts2 = tof_info(1.0e4, 1.0e4)
    assert np.float32(ts2) < np.float32(405758.0)
    lors = H.siddon_lors(50, seed=2)
    g, c, s = H.SID_GRID, H.SID_CENTER, H.SID_SIZE
    x = np.ones(g[::-1], np.float32)
    on = O.siddon_project(lors, x, g, c, s, tb, ts2, prec="f64")
    off = O.siddon_project(lors, x, g, c, s, tb, 405758.0, prec="f64")
    assert not np.allclose(on, off)


def test_fp32_port_close_to_fp64():
    """Off voxel faces the FP32 port (the reference's own arithmetic) agrees
    with the FP64 oracle to ~1e-6; on a face FP32 rounding picks the voxel."""
    lors = H.siddon_lors(2000, seed=13, special=False)
    g, c, s = H.SID_GRID, H.SID_CENTER, H.SID_SIZE
    x = H.image(g, seed=2)
    for ts2 in (400.0, 1.0e6):
        a = O.siddon_project(lors, x, g, c, s, 6.0, ts2, prec="f64")
        b = O.siddon_project(lors, x, g, c, s, 6.0, ts2, prec="f32")
        assert H.rel_l2(b, a) < 1e-5
