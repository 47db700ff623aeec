"""Efficiency map: backprojection of every crystal pair, then the reference's
finalisation (src/python/srf/preprocess/merge_map.py) on the device.

compute_tor_effmap   Maplors over all unordered crystal pairs of a scanner,
                     in chunks, through SplitLoRsModel.maplors (values 1).
crop_image           merge_map.py:47-65 (radial and axial crop).
finalize             merge_map.py:37-41: crop, divide by the max, invert where
                     > 1e-6 (the stored map is the inverse, used as x*E*bp).
initial_image        MasterLoader.create_image, data/loader.py:28-37.

The reference's map task, reproduced as well:
ring_maps            SRFApp._make_map_single_ring (srfapp_new.py:186-201): the
                     map of ring 0 with each ring.
merge_effmap         merge_map.py:8-44: those maps summed with the axial
                     translation of every ring pair, then finalize.
ordered_pair_map     the map task without translation (srfapp_new.py:153-164):
                     every ordered ring pair, so cross-ring pairs count twice.
merge_effmap_full    merge_map.py:68-88: crop, divide by the max, invert where
                     > 1e-7 (values at or below stay as they are).
compute_tor_effmap backprojects each unordered pair once instead.
"""
import torch

from .data import Image, split_from_groups
from .preprocess import map_process


def _meshes(grid, center, size, device):
    """RingBlock(size, grid, center, 0).get_meshes() (scanner/pet/block.py:77-108):
    np.meshgrid's 'xy' indexing orders points y-major, then x, then z."""
    axes = []
    for k in range(3):
        interval = size[k] / grid[k]
        start = center[k] - (size[k] - interval) / 2
        end = center[k] + (size[k] - interval) / 2
        axes.append(torch.linspace(start, end, grid[k], dtype=torch.float64, device=device))
    mx, my, mz = torch.meshgrid(axes[0], axes[1], axes[2], indexing="xy")
    return mx.reshape(-1), my.reshape(-1), mz.reshape(-1)


def crop_mask(grid, center, size, inner_radius, axial_length, crop_ratio=0.95, device="cpu"):
    """Boolean (nx, ny, nz) mask of voxels kept by crop_image."""
    mx, my, mz = _meshes(grid, center, size, device)
    xy = torch.sqrt(mx * mx + my * my)
    keep = ~((xy > crop_ratio * inner_radius) | (mz.abs() >= axial_length / 2))
    return keep.reshape(grid[0], grid[1], grid[2])


def crop_image(image, grid, center, size, inner_radius, axial_length, crop_ratio=0.95):
    m = crop_mask(grid, center, size, inner_radius, axial_length, crop_ratio, image.device)
    return torch.where(m, image, torch.zeros((), dtype=image.dtype, device=image.device))


def finalize(summed, grid, center, size, inner_radius, axial_length, crop_ratio=0.95):
    """Cropped, max-normalised, inverted efficiency map (float32)."""
    m = crop_image(summed.to(torch.float64), grid, center, size, inner_radius, axial_length,
                   crop_ratio)
    m = m / torch.max(m)
    big = m > 1e-6
    out = torch.where(big, 1.0 / torch.where(big, m, torch.ones_like(m)), torch.zeros_like(m))
    return out.to(torch.float32)


def initial_image(grid, center, size, inner_radius, axial_length, device, crop_ratio=0.95):
    ones = torch.ones(tuple(grid), dtype=torch.float32, device=device)
    return crop_image(ones, grid, center, size, inner_radius, axial_length, crop_ratio)


def compute_tor_effmap(scanner, model, grid, center, size, device, chunk=8_000_000,
                       max_pairs=None, tiled=False):
    """Sum over all unordered crystal pairs of Maplors (values 1), (nx, ny, nz).
    Pairs are enumerated on the device.  tiled=True bins each chunk's map LORs
    into image tiles and scatters through LDS tiles; for the single pass of a
    map the plan costs more than it saves (0.98 s against 0.77 s for the 75.5M
    pairs at 200x200x100), so the default is the direct scatter."""
    out = torch.zeros(tuple(grid), dtype=torch.float32, device=device)
    img = Image(out, center, size)
    total = scanner.nb_pairs() if max_pairs is None else min(max_pairs, scanner.nb_pairs())
    for start in range(0, total, chunk):
        stop = min(total, start + chunk)
        pairs = scanner.all_pairs_device((start, stop), device)
        groups = map_process(pairs)
        data = split_from_groups(groups, without_tof=True)
        if tiled:
            plans = model.make_plans(img, data, ncols=7)
            model.maplors(data, img, out=out, plans=plans)
            for p in plans.values():
                if p is not None:
                    p.close()
        else:
            model.maplors(data, img, out=out)
    return out


def _map_lors(model, lors, img, out):
    groups = map_process(lors)
    model.maplors(split_from_groups(groups, without_tof=True), img, out=out)


def ring_maps(scanner, model, grid, center, size, device, ring1=0):
    """[map of the LORs of ring ``ring1`` with ring ir, for ir in range(nb_rings)]
    (srfapp_new.py:186-201, each saved there as effmap_{ir}.npy), (nx, ny, nz)."""
    maps = []
    for ir in range(scanner.nb_rings):
        out = torch.zeros(tuple(grid), dtype=torch.float32, device=device)
        lors = scanner.ring_pair_lors(ring1, ir, device)
        if len(lors):
            _map_lors(model, lors, Image(out, center, size), out)
        maps.append(out)
    return maps


def merge_effmap(maps, grid, center, size, inner_radius, axial_length, crop_ratio=0.95):
    """merge_map.py:8-44 on the device: map ir (ring 0 with ring ir) is added
    at every axial shift jr = 0 .. nb_rings - ir - 1 (image layer jr onwards
    takes layers 0 .. L - jr - 1), then cropped, normalised and inverted
    (finalize).  Float64 sums, float32 result."""
    nz = maps[0].shape[2]
    nb_rings = len(maps)
    if nb_rings - 1 > nz:
        # the reference's slices final[jr:L] += temp[0:L - jr] no longer match
        raise ValueError(f"{nb_rings} rings need at least {nb_rings - 1} image layers, got {nz}")
    acc = torch.zeros(maps[0].shape, dtype=torch.float64, device=maps[0].device)
    for ir, m in enumerate(maps):
        m = m.to(torch.float64)
        for jr in range(min(nb_rings - ir, nz)):
            acc[:, :, jr:] += m[:, :, :nz - jr]
    return finalize(acc, grid, center, size, inner_radius, axial_length, crop_ratio)


def ordered_pair_map(scanner, model, grid, center, size, device):
    """Sum of the maps of every ordered ring pair (srfapp_new.py:153-164;
    merge_map.py:78-82 adds the saved maps), (nx, ny, nz) float32."""
    out = torch.zeros(tuple(grid), dtype=torch.float32, device=device)
    img = Image(out, center, size)
    for ir1 in range(scanner.nb_rings):
        for ir2 in range(scanner.nb_rings):
            lors = scanner.ring_pair_lors(ir1, ir2, device)
            if len(lors):
                _map_lors(model, lors, img, out)
    return out


def merge_effmap_full(summed, grid, center, size, inner_radius, axial_length, crop_ratio=0.95):
    """merge_map.py:84-88: crop, divide by the max, then 1 / value where
    value > 1e-7; values at or below 1e-7 are kept (not zeroed), as there."""
    m = crop_image(summed.to(torch.float64), grid, center, size, inner_radius, axial_length,
                   crop_ratio)
    m = m / torch.max(m)
    big = m > 1e-7
    out = torch.where(big, 1.0 / torch.where(big, m, torch.ones_like(m)), m)
    return out.to(torch.float32)
