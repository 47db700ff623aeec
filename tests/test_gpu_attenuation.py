"""Attenuation weights (Siddon line integrals of a mu-map) and the attenuated
efficiency map."""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

GRID, CENTER, SIZE = (100, 100, 28), (0.0, 0.0, 0.0), (400.0, 400.0, 112.0)


def test_line_integrals_vs_oracle_and_closed_form(cuda):
    from oracle import lib as O
    from srf_amd import attenuation as A
    from srf_amd.data import Image
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner()
    mu = A.water_cylinder(GRID, CENTER, SIZE, 150.0, 54.4, device=cuda)
    pairs = torch.from_numpy(sc.all_pairs((0, 200000))).to(cuda)
    li = A.line_integrals(Image(mu, CENTER, SIZE), pairs).cpu().numpy()
    lors = np.vstack([pairs.cpu().numpy().T, np.zeros((1, pairs.shape[0]), np.float32)])
    mut = np.ascontiguousarray(mu.cpu().numpy().transpose(2, 1, 0))  # reference frame [z][y][x]
    ref = O.siddon_project(lors, mut, GRID, CENTER, SIZE, 1.0, 1.0e6, prec="f64")
    L2 = np.sum((lors[3:6].astype(np.float64) - lors[0:3]) ** 2, axis=0)
    assert H.rel_l2(li, ref * L2) <= 1e-5
    # same-ring LOR through the axis: chord 300 mm of water
    p = sc.crystal_xyz(np.array([10, 10]), np.array([0, 192]))
    li2 = A.line_integrals(Image(mu, CENTER, SIZE),
                           torch.from_numpy(np.hstack([p[0], p[1]])[None].astype(np.float32))
                           .to(cuda)).item()
    assert li2 == pytest.approx(0.0096 * 300.0, rel=0.02)
    w = A.attenuation_weights(Image(mu, CENTER, SIZE), pairs[:5]).cpu().numpy()
    np.testing.assert_allclose(w, 1.0 / np.exp(-li[:5]), rtol=1e-6)


def test_attenuated_effmap(cuda):
    from srf_amd import attenuation as A, effmap, physics
    from srf_amd.data import Image
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner()
    model = physics.SplitLoRsModel(8.0, False)
    grid, size = (50, 50, 14), (400.0, 400.0, 112.0)
    zero = Image(torch.zeros(grid, device=cuda), CENTER, size)
    n = 300000
    plain = effmap.compute_tor_effmap(sc, model, grid, CENTER, size, cuda, max_pairs=n)
    att0 = A.attenuated_tor_effmap(sc, model, zero, grid, CENTER, size, cuda, max_pairs=n)
    assert H.rel_l2(att0.cpu().numpy(), plain.cpu().numpy()) <= 1e-6
    mu = Image(A.water_cylinder(grid, CENTER, size, 150.0, 54.4, device=cuda), CENTER, size)
    att = A.attenuated_tor_effmap(sc, model, mu, grid, CENTER, size, cuda, max_pairs=n)
    a, b = att.cpu().numpy(), plain.cpu().numpy()
    assert np.all(a <= b * (1 + 1e-5) + 1e-6)
    inner = b > 0.01 * b.max()
    ratio = (a[inner] / b[inner])
    assert ratio.min() > np.exp(-0.0096 * 300 * 1.05) and ratio.max() <= 1.0 + 1e-5


def test_tiled_effmap_equals_untiled(cuda):
    """Maplors over crystal pairs through tile plans (the default) against the
    untiled scatter, plain and attenuated."""
    from srf_amd import attenuation as A, effmap, physics
    from srf_amd.data import Image
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner()
    model = physics.SplitLoRsModel(4.0, False)
    grid, size = (100, 100, 50), (400.0, 400.0, 110.0)
    n = 400000
    t = effmap.compute_tor_effmap(sc, model, grid, CENTER, size, cuda, max_pairs=n, chunk=150000,
                                  tiled=True)
    u = effmap.compute_tor_effmap(sc, model, grid, CENTER, size, cuda, max_pairs=n, chunk=150000,
                                  tiled=False)
    assert float(u.abs().sum()) > 0
    assert H.rel_l2(t.cpu().numpy(), u.cpu().numpy()) <= 1e-6
    mu = Image(A.water_cylinder(grid, CENTER, size, 150.0, 54.4, device=cuda), CENTER, size)
    at = A.attenuated_tor_effmap(sc, model, mu, grid, CENTER, size, cuda, max_pairs=n, tiled=True)
    au = A.attenuated_tor_effmap(sc, model, mu, grid, CENTER, size, cuda, max_pairs=n, tiled=False)
    assert H.rel_l2(at.cpu().numpy(), au.cpu().numpy()) <= 1e-6
