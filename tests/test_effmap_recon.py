"""Efficiency-map finalisation, initial image, scanner geometry and the
OSEM subset rules, on the CPU (product torch code against the numpy oracle
restatement of merge_map.py / loader.py, and the reference tests' rules)."""
import numpy as np
import pytest
import torch

from oracle import effmap as OE
from srf_amd import effmap, recon
from srf_amd.scanner import CylindricalScanner


@pytest.mark.parametrize("grid,center,size", [
    ((20, 20, 10), (0.0, 0.0, 0.0), (800.0, 800.0, 220.0)),
    ((12, 16, 9), (3.0, -2.0, 1.0), (780.0, 800.0, 120.0)),
])
def test_crop_and_finalize_match_oracle(grid, center, size):
    rng = np.random.default_rng(0)
    s = rng.uniform(0, 5, grid)
    s[0, 0, 0] = 0.0
    ref_c = OE.crop_image(390.0, 108.8, s, grid, center, size, 0.95)
    got_c = effmap.crop_image(torch.from_numpy(s), grid, center, size, 390.0, 108.8)
    np.testing.assert_array_equal(got_c.numpy(), ref_c)
    ref = OE.finalize_effmap(s, 390.0, 108.8, grid, center, size)
    got = effmap.finalize(torch.from_numpy(s), grid, center, size, 390.0, 108.8)
    np.testing.assert_allclose(got.numpy(), ref.astype(np.float32), rtol=1e-7)


def test_finalize_closed_form():
    """max -> 1, values <= 1e-6 of the max -> 0, others inverted."""
    s = np.zeros((4, 4, 2))
    s[1, 1, 0] = 8.0
    s[2, 2, 1] = 2.0
    s[1, 2, 1] = 8e-7
    e = effmap.finalize(torch.from_numpy(s), (4, 4, 2), (0, 0, 0), (40, 40, 4), 390.0, 108.8)
    assert e[1, 1, 0] == 1.0 and e[2, 2, 1] == 4.0 and e[1, 2, 1] == 0.0
    assert e.sum().item() == 5.0


def test_initial_image_matches_oracle():
    grid, center, size = (20, 20, 10), (0, 0, 0), (800.0, 800.0, 220.0)
    ref = OE.initial_image(390.0, 108.8, grid, center, size)
    got = effmap.initial_image(grid, center, size, 390.0, 108.8, "cpu")
    np.testing.assert_array_equal(got.numpy(), ref)


def test_subset_split_reference_rule():
    """test_worker.py:105-121: 30 rows, 3 subsets -> rows [10 s, 10 s + 10)."""
    assert [recon.subset_bounds(30, s, 3) for s in range(3)] == [(0, 10), (10, 20), (20, 30)]
    # every row in exactly one subset for uneven sizes
    for n, k in ((31, 3), (7, 10), (1000, 7)):
        rows = np.concatenate([np.arange(*recon.subset_bounds(n, s, k)) for s in range(k)])
        np.testing.assert_array_equal(rows, np.arange(n))


def test_scanner_geometry():
    sc = CylindricalScanner()
    z = sc.ring_z()
    assert z[0] == pytest.approx(-52.7) and z[-1] == pytest.approx(52.7)
    assert sc.axial_length == pytest.approx(108.8)
    p = sc.crystal_xyz(np.array([0]), np.array([0]))[0]
    np.testing.assert_allclose(p, [400.0, 0.0, -52.7])
    p = sc.crystal_xyz(np.array([31]), np.array([96]))[0]  # 90 degrees
    np.testing.assert_allclose(p, [0.0, 400.0, 52.7], atol=1e-9)
    assert sc.nb_pairs() == 12288 * 12287 // 2


def test_all_pairs_enumeration():
    sc = CylindricalScanner(nb_rings=3, nb_crystals=8)
    nc = 24
    pairs = sc.all_pairs()
    assert pairs.shape == (nc * (nc - 1) // 2, 6)
    ref = []
    for a in range(nc):
        for b in range(a + 1, nc):
            ref.append(np.concatenate([sc.crystal_xyz(a // 8, a % 8), sc.crystal_xyz(b // 8, b % 8)]))
    np.testing.assert_allclose(pairs, np.array(ref, np.float32))
    # chunks reproduce the full enumeration
    parts = np.vstack([sc.all_pairs((s, min(s + 50, len(ref)))) for s in range(0, len(ref), 50)])
    np.testing.assert_array_equal(parts, pairs)


def test_all_pairs_device_matches_host():
    """The device enumeration (run here on the CPU device) gives the host
    enumeration's pairs, order and float32 values bit for bit."""
    import torch
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner(nb_rings=4, nb_crystals=48)
    n = sc.nb_pairs()
    for chunk in ((0, n), (17, 1000), (n - 300, n)):
        ref = sc.all_pairs(chunk)
        got = sc.all_pairs_device(chunk, torch.device("cpu")).numpy()
        np.testing.assert_array_equal(got, ref)
    big = CylindricalScanner()
    m = big.nb_pairs()
    for chunk in ((0, 2000), (m // 2, m // 2 + 3000), (m - 5000, m)):
        np.testing.assert_array_equal(big.all_pairs_device(chunk, torch.device("cpu")).numpy(),
                                      big.all_pairs(chunk))


def _delta_maps(grid, nb_rings, seed):
    """Ring maps that are single voxels: map ir holds value v_ir at (ix, iy, z0)."""
    rng = np.random.default_rng(seed)
    vals = rng.uniform(0.5, 2.0, nb_rings)
    maps = []
    for v in vals:
        m = np.zeros(grid)
        m[grid[0] // 2, grid[1] // 2, 1] = v
        maps.append(m)
    return maps, vals


def test_merge_effmap_translation_closed_form():
    """merge_map.py:18-25: layer z = 1 + jr collects v_ir for every ring ir with
    jr < nb_rings - ir, so the summed map there is sum_{ir < nb_rings - jr} v_ir;
    the oracle's numpy restatement and the device code agree with it."""
    grid, center, size = (8, 8, 12), (0.0, 0.0, 0.0), (40.0, 40.0, 36.0)
    nb_rings = 6
    maps, vals = _delta_maps(grid, nb_rings, 3)
    expect = np.zeros(grid)
    for jr in range(nb_rings):
        expect[4, 4, 1 + jr] = vals[:nb_rings - jr].sum()
    # the unnormalised sum, through the oracle with crop and inversion undone
    ref = OE.merge_effmap(maps, 390.0, 1000.0, grid, center, size)
    got = effmap.merge_effmap([torch.from_numpy(m) for m in maps], grid, center, size, 390.0,
                              1000.0).numpy()
    fin = OE.finalize_effmap(expect, 390.0, 1000.0, grid, center, size)
    np.testing.assert_allclose(ref, fin, rtol=1e-12)
    np.testing.assert_allclose(got, fin.astype(np.float32), rtol=1e-7)
    assert np.count_nonzero(got) == nb_rings


def test_merge_effmap_matches_oracle_random():
    grid, center, size = (10, 12, 9), (1.0, -2.0, 0.5), (780.0, 800.0, 30.6)
    rng = np.random.default_rng(5)
    maps = [rng.uniform(0, 3, grid) for _ in range(7)]
    ref = OE.merge_effmap(maps, 390.0, 30.6, grid, center, size)
    got = effmap.merge_effmap([torch.from_numpy(m) for m in maps], grid, center, size, 390.0,
                              30.6)
    np.testing.assert_allclose(got.numpy(), ref.astype(np.float32), rtol=1e-6)
    with pytest.raises(ValueError):
        effmap.merge_effmap([torch.from_numpy(m[:, :, :3]) for m in maps], grid, center, size,
                            390.0, 30.6)


def test_merge_effmap_full_keeps_small_values():
    """merge_map.py:86-87: only values > 1e-7 of the max are inverted; the rest
    stay as normalised (no zeroing, unlike merge_effmap)."""
    grid, center, size = (4, 4, 2), (0.0, 0.0, 0.0), (40.0, 40.0, 4.0)
    s = np.zeros(grid)
    s[1, 1, 0] = 8.0
    s[2, 2, 1] = 2.0
    s[1, 2, 1] = 4e-7  # 5e-8 of the max: kept as 5e-8
    got = effmap.merge_effmap_full(torch.from_numpy(s), grid, center, size, 390.0, 108.8)
    ref = OE.merge_effmap_full(s, 390.0, 108.8, grid, center, size)
    assert got[1, 1, 0] == 1.0 and got[2, 2, 1] == 4.0
    assert got[1, 2, 1].item() == pytest.approx(5e-8, rel=1e-6)
    np.testing.assert_allclose(got.numpy(), ref.astype(np.float32), rtol=1e-7)


def test_ring_pair_lors_rule():
    """pet.py:160-167 and block.py:140-152 on a small scanner: same ring
    i1 < i2, different rings i1 != i2, i1 outer; B(B-1)/2 and B(B-1) LORs."""
    sc = CylindricalScanner(nb_rings=3, nb_crystals=7)
    tab = sc.crystal_table()
    b = sc.nb_crystals
    for r1 in range(3):
        for r2 in range(3):
            got = sc.ring_pair_lors(r1, r2).numpy()
            pairs = [(i1, i2) for i1 in range(b) for i2 in range(b)
                     if (i1 < i2 if r1 == r2 else i1 != i2)]
            assert len(got) == (b * (b - 1) // 2 if r1 == r2 else b * (b - 1))
            want = np.array([np.concatenate([tab[r1 * b + i1], tab[r2 * b + i2]])
                             for i1, i2 in pairs], dtype=np.float32)
            np.testing.assert_array_equal(got, want)
