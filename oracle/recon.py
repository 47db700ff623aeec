"""TEST INFRASTRUCTURE ONLY — the reference's reconstruction semantics on
the CPU oracle.

group_frame / back_to_world   split_lors.py:54-70, 79-85, 112: the image is
                              transposed by perm(a), the op sees grid/center/
                              size of the transposed image reversed, and the
                              result is transposed back with perm_back(a).
project_split / backproject_split  the SplitLoRsModel projection /
                              backprojection (sum of the groups).
osem                          ReconStep + mlem_update (recon_step.py:33-49)
                              over contiguous subsets (worker.py:79-85), or
                              PSFReconStep (recon_step.py:56-115) with psf.
siddon_osem                   the same over CompleteLoRsModel (Siddon,
                              complete_lors.py:59-98).
"""
import numpy as np

from . import lib as O

PERM = {"z": [2, 1, 0], "y": [1, 2, 0], "x": [0, 2, 1]}
PERM_BACK = {"z": [2, 1, 0], "y": [2, 0, 1], "x": [0, 2, 1]}


def group_frame(x, center, size, axis):
    p = PERM[axis]
    xa = np.ascontiguousarray(np.transpose(x, p))
    grid = list(xa.shape[::-1])
    cen = [center[k] for k in p][::-1]
    siz = [size[k] for k in p][::-1]
    return xa, grid, cen, siz


def back_to_world(img_op, axis):
    """img_op is laid out [gz][gy][gx] = the transposed image's shape."""
    return np.transpose(img_op, PERM_BACK[axis])


def project_split(groups, x, center, size, kw, geo, tof_bin, tof_sigma2, prec="f64"):
    out = {}
    for a in ("x", "y", "z"):
        xa, g, c, s = group_frame(x, center, size, a)
        lors = groups[a]
        if lors.shape[1] == 0:
            out[a] = np.zeros(0)
            continue
        out[a] = O.tor_project(lors, xa, g, c, s, kw, geo, tof_bin, tof_sigma2, prec=prec)
    return out


def backproject_split(groups, values, x_shape, center, size, kw, geo, tof_bin, tof_sigma2,
                      prec="f64"):
    acc = np.zeros(x_shape, np.float64)
    dummy = np.zeros(x_shape, np.float32)
    for a in ("x", "y", "z"):
        lors = groups[a]
        if lors.shape[1] == 0:
            continue
        _, g, c, s = group_frame(dummy, center, size, a)
        b = O.tor_backproject(lors, values[a], g, c, s, kw, geo, tof_bin, tof_sigma2, prec=prec)
        acc += back_to_world(b, a)
    return acc


def map_split(groups7, x_shape, center, size, kw, geo, prec="f64"):
    """Sum over groups of Maplors (values 1) of map rows [7, N_a] per group,
    in the world (nx, ny, nz) frame."""
    acc = np.zeros(x_shape, np.float64)
    dummy = np.zeros(x_shape, np.float32)
    for a in ("x", "y", "z"):
        lors = groups7[a]
        if lors.shape[1] == 0:
            continue
        _, g, c, s = group_frame(dummy, center, size, a)
        m = O.tor_map(lors, np.ones(lors.shape[1], np.float32), g, c, s, kw, geo, prec=prec)
        acc += back_to_world(m, a)
    return acc


def subset_bounds(n, k, K):
    return (k * n) // K, ((k + 1) * n) // K


def osem(groups, x0, emap, center, size, kw, geo, tof_bin, tof_sigma2, nb_iterations,
         nb_subsets, prec="f64", psf=None):
    """x <- x * E * sum_groups A^T(1 / A x) per subset (the list-mode ratio of
    the reference, with its 1e-7 threshold inside the backprojection).  With
    psf = (psf_xy, psf_z): PSFReconStep (recon_step.py:64-115), the projector
    sees PSF(x) and the backprojection is blurred by PSF^T before the update."""
    from . import psf as OPSF
    x = np.array(x0, np.float64)
    for _ in range(nb_iterations):
        for k in range(nb_subsets):
            sub = {}
            for a in ("x", "y", "z"):
                lo, hi = subset_bounds(groups[a].shape[1], k, nb_subsets)
                sub[a] = np.ascontiguousarray(groups[a][:, lo:hi])
            seen = x if psf is None else OPSF.psf_forward(x, *psf)
            p = project_split(sub, seen.astype(np.float32), center, size, kw, geo, tof_bin,
                              tof_sigma2, prec)
            v = {a: p[a].astype(np.float32) for a in p}
            bp = backproject_split(sub, v, x.shape, center, size, kw, geo, tof_bin, tof_sigma2,
                                   prec)
            if psf is not None:
                bp = OPSF.psf_backward(bp, *psf)
            x = x * emap * bp
    return x


def siddon_osem(lors, x0, emap, center, size, tof_bin, tof_sigma2, nb_iterations, nb_subsets,
                weights=None, prec="f64"):
    """CompleteLoRsModel + ReconStep + mlem_update (complete_lors.py:59-98,
    recon_step.py:33-49) over contiguous subsets of the [7, N] Siddon rows.
    The op sees the world image (nx, ny, nz) transposed [2, 1, 0], i.e. laid
    out [gz][gy][gx] with grid (nx, ny, nz) and the world center and size
    (complete_lors.py:61-70 transposes the image and reverses the reversed
    geometry).  weights (counts per LOR, optional): lors_value = p / w, so the
    backprojection adds w_i a_ij / p_i; without them lors_value = p."""
    x = np.array(x0, np.float64)
    grid = list(x.shape)
    n = lors.shape[1]
    for _ in range(nb_iterations):
        for k in range(nb_subsets):
            lo, hi = subset_bounds(n, k, nb_subsets)
            sub = np.ascontiguousarray(lors[:, lo:hi])
            xt = np.ascontiguousarray(np.transpose(x, (2, 1, 0))).astype(np.float32)
            p = O.siddon_project(sub, xt, grid, center, size, tof_bin, tof_sigma2, prec=prec)
            v = p if weights is None else p / np.asarray(weights[lo:hi], np.float64)
            bp = O.siddon_backproject(sub, v.astype(np.float32), grid, center, size, tof_bin,
                                      tof_sigma2, prec=prec)
            x = x * emap * np.transpose(bp, (2, 1, 0))
    return x
