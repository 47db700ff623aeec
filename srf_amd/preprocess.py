"""LOR preprocessing for the TOR model, on the device.

Restates src/python/srf/preprocess/function/on_tor_lors.py (reference) with
torch float64 tensor ops in the same order as the reference's numpy
expressions, so results are identical on any device.  Output groups are
structure-of-arrays float32 tensors in the group frame, ready for the ops:

  recon_process(lors[N, 7], tof_res)  -> {'x','y','z'}: [10, N_a]
      rows p1 (3), p2 (3), TOF centre (3), sigma2 factor, in the group frame
  map_process(lors[N, 6])             -> {'x','y','z'}: [7, N_a]
      rows p1 (3), p2 (3), sigma2 factor

Row order inside a group is the reference's: ascending event index
(np.intersect1d) for rows with d_axis >= 0, then those with d_axis < 0
(SwapPointsOrder).  ``order`` maps output rows to input rows.
"""
import torch

AXES = ("x", "y", "z")
GAUSSIAN_FACTOR = 2.35482005  # on_tor_lors.py:220
# Column permutations into the group frame, on_tor_lors.py:246-247, 259-260.
PERM_RECON = {"x": [1, 2, 0, 4, 5, 3, 7, 8, 6], "y": [0, 2, 1, 3, 5, 4, 6, 8, 7]}
PERM_MAP = {"x": [1, 2, 0, 4, 5, 3], "y": [0, 2, 1, 3, 5, 4]}


def sigma2_factor(lors):
    """on_tor_lors.py:137-166."""
    dx = lors[:, 3] - lors[:, 0]
    dy = lors[:, 4] - lors[:, 1]
    dz = lors[:, 5] - lors[:, 2]
    p_dis2 = (dx * dx + dy * dy) + dz * dz
    dsq = dx * dx + dy * dy
    rsq = lors[:, 0] * lors[:, 0] + lors[:, 3] * lors[:, 3] + lors[:, 1] * lors[:, 1] + \
        lors[:, 4] * lors[:, 4]
    return dsq * dsq / rsq / p_dis2 / 2


def dominant_masks(lors):
    """Partition rule, on_tor_lors.py:70-77."""
    d = (lors[:, 3:6] - lors[:, 0:3]).abs()
    ax, ay, az = d[:, 0], d[:, 1], d[:, 2]
    return {"x": (ax >= ay) & (ax >= az), "y": (ay > ax) & (ay >= az),
            "z": (az > ax) & (az > ay)}


def _partition_swap(lors, axis, mask):
    """Partition then SwapPointsOrder (on_tor_lors.py:175-193)."""
    a = AXES.index(axis)
    idx = torch.nonzero(mask, as_tuple=False).reshape(-1)
    part = lors[idx]
    d = part[:, 3 + a] - part[:, a]
    pos = d >= 0
    neg = d < 0
    positive = part[pos]
    negative = part[neg]
    if lors.shape[1] > 7:
        swapped = torch.cat([negative[:, 3:6], negative[:, 0:3], 0 - negative[:, 6:7],
                             negative[:, 7:]], 1)
    else:
        swapped = torch.cat([negative[:, 3:6], negative[:, 0:3], negative[:, 6:7]], 1)
    order = torch.cat([idx[pos], idx[neg]])
    return torch.cat([positive, swapped], 0), order


def cut_lors(lors, tof_res):
    """CutLoRs, on_tor_lors.py:196-229: output (p0, p1, centre, extra)."""
    p0 = lors[:, 0:3]
    p_diff = lors[:, 3:6] - p0
    p_dis = torch.sqrt((p_diff[:, 0] * p_diff[:, 0] + p_diff[:, 1] * p_diff[:, 1]) +
                       p_diff[:, 2] * p_diff[:, 2])
    dcos = p_diff / p_dis[:, None]
    ratio = 0.5 - (lors[:, 6] / p_dis)
    center = ratio[:, None] * p_diff + p0
    c0 = center - p0
    dcs = torch.sqrt((c0[:, 0] * c0[:, 0] + c0[:, 1] * c0[:, 1]) + c0[:, 2] * c0[:, 2])
    q0 = p0.clone()
    q1 = lors[:, 3:6].clone()
    if tof_res < 10000:
        limit = tof_res * 0.15 / GAUSSIAN_FACTOR * 3
        index = dcs > limit
        q0 = torch.where(index[:, None], center - limit * dcos, q0)
        q1 = torch.where(index[:, None], center + limit * dcos, q1)
    return torch.cat([q0, q1, center, lors[:, 7:]], 1)


def _as64(lors, ncols):
    if not isinstance(lors, torch.Tensor):
        lors = torch.as_tensor(lors)
    if lors.dim() != 2 or lors.shape[1] != ncols:
        raise ValueError(f"expected [N, {ncols}] rows, got {tuple(lors.shape)}")
    return lors.to(torch.float64)


def recon_process(lors, tof_res, return_order=False):
    """on_tor_lors.py:251-262.  lors: [N, 7] rows (x1, y1, z1, x2, y2, z2, tof)."""
    lors = _as64(lors, 7)
    lors = torch.cat([lors, sigma2_factor(lors)[:, None]], 1)
    masks = dominant_masks(lors)
    out, orders = {}, {}
    for a in AXES:
        g, order = _partition_swap(lors, a, masks[a])
        g = cut_lors(g, tof_res)
        if a in PERM_RECON:
            g = torch.cat([g[:, PERM_RECON[a]], g[:, 9:]], 1)
        out[a] = g.to(torch.float32).t().contiguous()
        orders[a] = order
    return (out, orders) if return_order else out


def map_process(lors, return_order=False):
    """on_tor_lors.py:243-248.  lors: [N, 6] rows (x1, y1, z1, x2, y2, z2)."""
    lors = _as64(lors, 6)
    lors = torch.cat([lors, sigma2_factor(lors)[:, None]], 1)
    masks = dominant_masks(lors)
    out, orders = {}, {}
    for a in AXES:
        g, order = _partition_swap(lors, a, masks[a])
        if a in PERM_MAP:
            g = torch.cat([g[:, PERM_MAP[a]], g[:, 6:]], 1)
        out[a] = g.to(torch.float32).t().contiguous()
        orders[a] = order
    return (out, orders) if return_order else out
