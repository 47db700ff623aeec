"""Efficiency map: backprojection of every crystal pair, then the reference's
finalisation (src/python/srf/preprocess/merge_map.py) on the device.

compute_tor_effmap   Maplors over all unordered crystal pairs of a scanner,
                     in chunks, through SplitLoRsModel.maplors (values 1).
crop_image           merge_map.py:47-65 (radial and axial crop).
finalize             merge_map.py:37-41: crop, divide by the max, invert where
                     > 1e-6 (the stored map is the inverse, used as x*E*bp).
initial_image        MasterLoader.create_image, data/loader.py:28-37.

The reference sums per-ring maps shifted along z (merge_map.py:18-25) and its
map task walks ordered ring pairs (srfapp_new.py:153-164), so cross-ring pairs
count twice; here each unordered pair is backprojected once.
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
