"""The reference's map task on the GPU: per-ring maps merged with the axial
translation (srfapp_new.py:186-201, merge_map.py:8-44) against the oracle
(oracle preprocessing + FP64 Maplors per ring pair + numpy merge), and the sum
over ordered ring pairs (srfapp_new.py:153-164) against its decomposition
into unordered pairs."""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

# 6 rings of 48 crystals on the standard radius; 3.4 mm ring pitch and
# 3.4 mm image layers, so one ring shift is one image layer
GRID, CENTER, SIZE = (24, 24, 8), (0.0, 0.0, 0.0), (480.0, 480.0, 27.2)
KW = 30.0


def _scanner():
    from srf_amd.scanner import CylindricalScanner
    return CylindricalScanner(nb_rings=6, nb_crystals=48)


def test_ring_maps_merge_vs_oracle(cuda):
    from oracle import effmap as OE, preprocess as OP, recon as OR
    from srf_amd import effmap, physics
    sc = _scanner()
    model = physics.SplitLoRsModel(KW, True)
    maps = effmap.ring_maps(sc, model, GRID, CENTER, SIZE, cuda)
    ref_maps = []
    for ir in range(sc.nb_rings):
        lors = sc.ring_pair_lors(0, ir).numpy().astype(np.float64)
        g7 = {a: np.ascontiguousarray(g.T.astype(np.float32))
              for a, g in OP.map_process(lors).items()}
        ref_maps.append(OR.map_split(g7, GRID, CENTER, SIZE, KW, True))
        e = H.rel_l2(maps[ir].cpu().numpy(), ref_maps[-1])
        print(f"ring map 0-{ir}: rel_l2 vs f64 {e:.2e}")
        assert e <= 1e-5
    got = effmap.merge_effmap(maps, GRID, CENTER, SIZE, sc.inner_radius, sc.axial_length)
    ref = OE.merge_effmap(ref_maps, sc.inner_radius, sc.axial_length, GRID, CENTER, SIZE)
    e = H.rel_l2(got.cpu().numpy(), ref)
    print(f"merged translation map: rel_l2 vs f64 {e:.2e}")
    assert e <= 1e-5


def test_ordered_pair_map_decomposition(cuda):
    """Ordered ring pairs = every unordered crystal pair once, plus the
    cross-ring pairs with different crystal indices again, minus the
    cross-ring pairs with equal indices (pet.py:160-167 excludes them)."""
    from srf_amd import effmap, physics
    from srf_amd.data import Image
    sc = _scanner()
    model = physics.SplitLoRsModel(KW, True)
    full = effmap.ordered_pair_map(sc, model, GRID, CENTER, SIZE, cuda)
    unordered = effmap.compute_tor_effmap(sc, model, GRID, CENTER, SIZE, cuda)
    b = sc.nb_crystals
    extra = torch.zeros(GRID, dtype=torch.float32, device=cuda)
    axial = torch.zeros(GRID, dtype=torch.float32, device=cuda)
    tab = torch.from_numpy(sc.crystal_table()).to(cuda)
    for r1 in range(sc.nb_rings):
        for r2 in range(r1 + 1, sc.nb_rings):
            extra_lors = sc.ring_pair_lors(r1, r2, cuda)
            effmap._map_lors(model, extra_lors, Image(extra, CENTER, SIZE), extra)
            i = torch.arange(b, device=cuda)
            same = torch.cat([tab[r1 * b + i], tab[r2 * b + i]], dim=1)
            effmap._map_lors(model, same, Image(axial, CENTER, SIZE), axial)
    want = unordered.double() + extra.double() - axial.double()
    e = H.rel_l2(full.cpu().numpy(), want.cpu().numpy())
    print(f"ordered pairs vs decomposition: rel_l2 {e:.2e}")
    assert e <= 1e-6
