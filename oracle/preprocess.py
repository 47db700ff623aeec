"""TEST INFRASTRUCTURE ONLY — numpy restatement of the reference's LOR
preprocessing, src/python/srf/preprocess/function/on_tor_lors.py.

Input rows are (x1, y1, z1, x2, y2, z2, tof) in float64 (recon) or
(x1..z2) (map).  Arithmetic follows the reference's numpy expressions in the
same order; group row order follows np.intersect1d (ascending index) and then
the positive-then-negative stacking of SwapPointsOrder.
"""
import numpy as np

AXES = ("x", "y", "z")
GAUSSIAN_FACTOR = 2.35482005  # on_tor_lors.py:220


def direction(lors):
    """on_tor_lors.py:54-56"""
    return lors[:, 3:6] - lors[:, 0:3]


def sigma2_factor(lors):
    """on_tor_lors.py:137-166: (dx^2+dy^2)^2 / (x1^2+x2^2+y1^2+y2^2) / |d|^2 / 2."""
    p_start = lors[:, 0:3]
    p_end = lors[:, 3:6]
    x_start, x_end = p_start[:, 0], p_end[:, 0]
    y_start, y_end = p_start[:, 1], p_end[:, 1]
    p_diff = p_end - p_start
    p_dis2 = np.sum(np.square(p_diff), 1)
    dsq = p_diff[:, 0] ** 2 + p_diff[:, 1] ** 2
    rsq = x_start ** 2 + x_end ** 2 + y_start ** 2 + y_end ** 2
    return dsq ** 2 / rsq / p_dis2 / 2


def dominent_by(axis, x_d, y_d, z_d):
    """on_tor_lors.py:70-77; returns ascending row indices."""
    if axis == "x":
        conds = [np.where(x_d >= y_d), np.where(x_d >= z_d)]
    elif axis == "y":
        conds = [np.where(y_d > x_d), np.where(y_d >= z_d)]
    else:
        conds = [np.where(z_d > x_d), np.where(z_d > y_d)]
    return np.intersect1d(*conds)


def partition(axis, lors):
    """Partition.kernel, on_tor_lors.py:97-107."""
    diff = np.abs(direction(lors))
    idx = dominent_by(axis, diff[:, 0], diff[:, 1], diff[:, 2])
    return lors[idx].reshape((-1, lors.shape[1])), idx


def swap_points_order(axis, lors):
    """SwapPointsOrder.kernel, on_tor_lors.py:175-193.  Rows with
    d_axis >= 0 first; the others after them with the end points swapped and,
    when there are more than 7 columns, the tof column negated."""
    a = AXES.index(axis)
    d = direction(lors)[:, a]
    positive = lors[d >= 0]
    negative = lors[d < 0]
    parts = [negative[:, 3:6], negative[:, 0:3]]
    if lors.shape[1] > 7:
        parts += [0 - np.expand_dims(negative[:, 6], 1), negative[:, 7:]]
    else:
        parts += [np.expand_dims(negative[:, 6], 1)]
    return np.vstack((positive, np.hstack(parts))), d >= 0


def cut_lors(tof_res, lors):
    """CutLoRs.kernel, on_tor_lors.py:196-229.  Output columns
    (p0, p1, lor_center, extra); only rows with dcs > limit are cut, on both
    sides."""
    p_diff = direction(lors)
    p_dis = np.sqrt(np.sum(np.square(p_diff), 1))
    dcos = np.array([p_diff[:, a] / p_dis for a in range(3)]).T.reshape((-1, 3))
    ratio = np.array(0.5 - (lors[:, 6] / p_dis))
    p0 = lors[:, 0:3].copy()
    p1 = lors[:, 3:6].copy()
    lor_center = np.array([ratio * p_diff[:, a] for a in range(3)]).T + lors[:, 0:3]
    dcs = np.sqrt(np.sum(np.square(lor_center - lors[:, 0:3]), 1))
    if tof_res < 10000:
        limit = tof_res * 0.15 / GAUSSIAN_FACTOR * 3
        index = dcs > limit
        p0[index] = lor_center[index] - limit * dcos[index, :]
        p1[index] = lor_center[index] + limit * dcos[index, :]
    return np.hstack((p0, p1, lor_center, lors[:, 7:]))


PERM_RECON = {"x": [1, 2, 0, 4, 5, 3, 7, 8, 6], "y": [0, 2, 1, 3, 5, 4, 6, 8, 7]}  # 259-260
PERM_MAP = {"x": [1, 2, 0, 4, 5, 3], "y": [0, 2, 1, 3, 5, 4]}  # 246-247


def group_order(lors):
    """Original row index of every output row, per group (the composition of
    Partition and SwapPointsOrder)."""
    out = {}
    for a in AXES:
        part, idx = partition(a, lors)
        d = direction(part)[:, AXES.index(a)]
        out[a] = np.concatenate([idx[d >= 0], idx[d < 0]])
    return out


def map_process(lors):
    """on_tor_lors.py:243-248: rows (x1..z2) -> per group [N_a, 7]
    (x1..z2 in the group frame, sigma2 factor)."""
    lors = np.hstack([lors, np.expand_dims(sigma2_factor(lors), 1)])
    out = {}
    for a in AXES:
        part, _ = partition(a, lors)
        g, _ = swap_points_order(a, part)
        if a in PERM_MAP:
            g[:, 0:6] = g[:, PERM_MAP[a]]
        out[a] = g
    return out


def recon_process(lors, tof_res):
    """on_tor_lors.py:251-262: rows (x1..z2, tof) -> per group [N_a, 10]
    (p1, p2, lor centre in the group frame, sigma2 factor)."""
    lors = np.hstack([lors, np.expand_dims(sigma2_factor(lors), 1)])
    out = {}
    for a in AXES:
        part, _ = partition(a, lors)
        g, _ = swap_points_order(a, part)
        g = cut_lors(tof_res, g)
        if a in PERM_RECON:
            g[:, 0:9] = g[:, PERM_RECON[a]]
        out[a] = g
    return out


def tof_info(tof_res_ps, tof_bin_ps):
    """SRFApp._get_tof_info, srfapp_new.py:244-251: (tof_bin [mm], tof_sigma2 [mm^2])."""
    limit = tof_res_ps * 0.15 / GAUSSIAN_FACTOR * 3
    tof_sigma2 = (limit ** 2) / 9
    return tof_bin_ps * 0.15, tof_sigma2
