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


def _spread8(v):
    """Spread the low 8 bits of v so that bit i moves to bit 4 i (for a
    4-way Morton interleave)."""
    v = v & 0xFF
    out = torch.zeros_like(v)
    for i in range(8):
        out |= ((v >> i) & 1) << (4 * i)
    return out


def locality_key(soa, center_z, zc_bin=40.0, slope_bin=0.04, pos_bin=6.0):
    """int64 key per event of a [10, N] group (op frame) that places events
    with nearby, nearly parallel tubes and nearby TOF centres next to each
    other: coarse TOF-centre position along the slice axis first, then a
    Morton interleave of the two slopes and the crossing point with the
    plane z = center_z."""
    x1, y1, z1 = soa[0].double(), soa[1].double(), soa[2].double()
    dx = soa[3].double() - x1
    dy = soa[4].double() - y1
    dz = soa[5].double() - z1
    dz = torch.where(dz == 0, torch.full_like(dz, 1e-12), dz)
    sx, sy = dx / dz, dy / dz
    X0 = x1 + sx * (center_z - z1)
    Y0 = y1 + sy * (center_z - z1)

    def q(v, step):
        return torch.clamp(torch.floor(v / step) + 128, 0, 255).long()

    zc = torch.clamp(torch.floor(soa[8].double() / zc_bin) + 64, 0, 127).long()
    m = _spread8(q(sx, slope_bin)) | (_spread8(q(sy, slope_bin)) << 1) | \
        (_spread8(q(X0, pos_bin)) << 2) | (_spread8(q(Y0, pos_bin)) << 3)
    return (zc << 32) | m


def locality_order(soa, nb_subsets, center_z=0.0):
    """Permutation of a group's events that sorts each OSEM subset's rows
    [k N / K, (k+1) N / K) by locality_key and leaves subset membership
    unchanged."""
    n = soa.shape[1]
    key = locality_key(soa, center_z)
    order = torch.empty(n, dtype=torch.int64, device=soa.device)
    for k in range(nb_subsets):
        lo, hi = (k * n) // nb_subsets, ((k + 1) * n) // nb_subsets
        if hi > lo:
            order[lo:hi] = lo + torch.argsort(key[lo:hi], stable=True)
    return order


def reorder_groups(groups, nb_subsets, centers=None):
    """Apply locality_order to every group; returns (groups, orders) with
    orders[a][i] = original row of output row i."""
    out, orders = {}, {}
    for a, soa in groups.items():
        cz = 0.0 if centers is None else centers[a]
        o = locality_order(soa, nb_subsets, cz)
        out[a] = soa[:, o].contiguous()
        orders[a] = o
    return out, orders
