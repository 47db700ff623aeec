"""Attenuation correction weights and the attenuated efficiency map.

The reference projects a mu-map with the Siddon model and turns each LOR's
line integral into a weight (graph/reconstruction/attenuation.py:10-43;
examples/attenuation/test_uniform_water_phantom.py:49-57):

    value  = Siddon projection of mu   (= sum mu l / L^2)
    weight = 1 / (exp(-value * L^2) + DBL_MIN)

Used as ``lors_value`` of Maplors (which divides by it), the weight turns the
efficiency map into sum_i w_ij exp(-int mu dl_i), the attenuated sensitivity.
"""
import sys

import torch

from .data import Image, ListModeData, split_from_groups
from .physics import CompleteLoRsModel
from .preprocess import map_process

DBL_MIN = sys.float_info.min


def line_integrals(mu, lors6):
    """int mu dl for every LOR (rows x1..z2, torch [N, 6] on the device) through
    the mu-map ``mu`` (an Image): the Siddon projection (non-TOF) times L^2."""
    n = lors6.shape[0]
    soa = torch.zeros(7, n, dtype=torch.float32, device=lors6.device)
    soa[0:6] = lors6.t().to(torch.float32)
    model = CompleteLoRsModel("attenuation", tof_sigma2=1.0e6, tof_bin=1.0)
    p = model.projection(mu, ListModeData(soa, soa=True)).values.double()
    d = (soa[3:6].double() - soa[0:3].double())
    L2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
    return p * L2


def attenuation_weights(mu, lors6):
    """1 / (exp(-int mu) + DBL_MIN) per LOR (float32)."""
    return (1.0 / (torch.exp(-line_integrals(mu, lors6)) + DBL_MIN)).to(torch.float32)


def water_cylinder(grid, center, size, radius, half_height, mu=0.0096, device="cpu"):
    """mu-map (nx, ny, nz) of a uniform cylinder, voxel centres inside."""
    axes = []
    for k in range(3):
        v = size[k] / grid[k]
        axes.append(center[k] - size[k] / 2 + (torch.arange(grid[k], dtype=torch.float64,
                                                            device=device) + 0.5) * v)
    X, Y, Z = torch.meshgrid(*axes, indexing="ij")
    inside = (X * X + Y * Y <= radius * radius) & (Z.abs() <= half_height)
    return (inside.to(torch.float32) * mu).contiguous()


def attenuated_tor_effmap(scanner, model, mu, grid, center, size, device, chunk=4_000_000,
                          max_pairs=None, tiled=False):
    """Sum over unordered crystal pairs of Maplors with the attenuation
    weights as lors_value: sum_i w_ij exp(-int mu dl_i).  Pairs are
    enumerated on the device; tiled=True scatters through tile plans (see
    effmap.compute_tor_effmap for the default)."""
    out = torch.zeros(tuple(grid), dtype=torch.float32, device=device)
    img = Image(out, center, size)
    total = scanner.nb_pairs() if max_pairs is None else min(max_pairs, scanner.nb_pairs())
    for start in range(0, total, chunk):
        stop = min(total, start + chunk)
        pairs = scanner.all_pairs_device((start, stop), device)
        w = attenuation_weights(mu, pairs)
        groups, orders = map_process(pairs, return_order=True)
        data = split_from_groups(groups, values={a: w[orders[a]] for a in orders},
                                 without_tof=True)
        if tiled:
            plans = model.make_plans(img, data, ncols=7)
            model.maplors(data, img, out=out, plans=plans)
            for p in plans.values():
                if p is not None:
                    p.close()
        else:
            model.maplors(data, img, out=out)
    return out
