"""PSF blur kernel (srfx_psf_apply) against the FP64 oracle restatement of
recon_step.py:79-110, and OSEM with the PSF step against the oracle OSEM."""
import numpy as np
import pytest
import torch
from scipy import sparse

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _mats(grid, seed, nnz_row=9, empty_rows=0, band=None):
    rng = np.random.default_rng(seed)
    nxy, nz = grid[0] * grid[1], grid[2]
    rows = np.repeat(np.arange(nxy), nnz_row)
    cols = rng.integers(0, nxy, rows.size)
    vals = rng.uniform(0.1, 1.0, rows.size)
    keep = rows >= empty_rows  # the first rows stay empty
    xy = sparse.csr_matrix((vals[keep], (rows[keep], cols[keep])), shape=(nxy, nxy))
    z = rng.uniform(0.0, 1.0, (nz, nz))
    if band is not None:  # zero outside z' - z in [-band[0], band[1]]
        d = np.arange(nz)[:, None] - np.arange(nz)[None, :]
        z[(d < -band[0]) | (d > band[1])] = 0.0
    return xy, z


@pytest.mark.parametrize("grid,empty,band", [((20, 24, 13), 0, None), ((7, 5, 1), 3, None),
                                             ((6, 9, 130), 5, None), ((3, 3, 64), 0, None),
                                             ((8, 6, 70), 0, (2, 5)), ((5, 5, 9), 2, (0, 3))])
def test_psf_vs_oracle(cuda, grid, empty, band):
    from oracle import psf as OPSF
    from srf_amd.data import Image
    from srf_amd.psf import PSF
    xy, z = _mats(grid, 1, empty_rows=empty, band=band)
    x = np.random.default_rng(2).uniform(0.5, 2.0, grid).astype(np.float32)
    p = PSF(xy, z, grid, cuda)
    img = Image(torch.from_numpy(x).to(cuda), (0, 0, 0), (1, 1, 1))
    f = p.forward(img).data.cpu().numpy()
    b = p.backward(img).data.cpu().numpy()
    ef = H.rel_l2(f, OPSF.psf_forward(x, xy, z))
    eb = H.rel_l2(b, OPSF.psf_backward(x, xy, z))
    print(f"psf {grid}: forward rel_l2 {ef:.2e}, backward {eb:.2e}")
    assert ef <= 1e-6 and eb <= 1e-6
    if empty:
        assert np.all(f.reshape(-1, grid[2])[:empty] == 0)


def test_psf_errors(cuda):
    from srf_amd import _lib
    from srf_amd.data import Image
    from srf_amd.psf import PSF
    grid = (4, 4, 3)
    xy, z = _mats(grid, 3)
    p = PSF(xy, z, grid, cuda)
    x = torch.ones(grid, device=cuda)
    with pytest.raises(_lib.SrfxError):
        p.forward(Image(x, (0, 0, 0), (1, 1, 1)), out=x)  # in place
    with pytest.raises(ValueError):
        p.forward(Image(torch.ones((4, 4, 4), device=cuda), (0, 0, 0), (1, 1, 1)))
    big = (1, 2, 2049)
    pb = PSF(sparse.eye(2), np.eye(2049), big, cuda)
    with pytest.raises(_lib.SrfxError):
        pb.forward(Image(torch.ones(big, device=cuda), (0, 0, 0), (1, 1, 1)))


def test_osem_with_psf_vs_oracle(cuda):
    from oracle import preprocess as OP, recon as OR
    from srf_amd import physics, preprocess, synth
    from srf_amd.data import Image, split_from_groups
    from srf_amd.psf import gaussian_psf
    from srf_amd.recon import OSEM
    from srf_amd.scanner import CylindricalScanner
    grid, center, size, kw = (40, 40, 24), (0.0, 0.0, 0.0), (400.0, 400.0, 132.0), 20.0
    sc = CylindricalScanner()
    ph = synth.c2_phantom(seed=5, scanner=sc)
    ev, _ = synth.tof_events(20000, ph, seed=9, scanner=sc, device="cpu")
    ev = ev.numpy()
    tb, ts2 = sc.tof_info()
    model = physics.SplitLoRsModel(kw, True, tof_sigma2=ts2, tof_bin=tb)
    groups = preprocess.recon_process(torch.from_numpy(ev).to(cuda), sc.tof_res_ps)
    E = np.random.default_rng(1).uniform(0.5, 1.5, grid).astype(np.float32)
    psf = gaussian_psf(grid, size, 25.0, 12.0, cuda)
    x = torch.ones(grid, device=cuda)
    osem = OSEM(model, split_from_groups(groups), Image(torch.from_numpy(E).to(cuda), center, size),
                nb_subsets=3, fused=True, tiled=True, psf=psf)
    osem.iterate(Image(x, center, size), 2)
    osem.close()
    ref_groups = {a: np.ascontiguousarray(g.T.astype(np.float32))
                  for a, g in OP.recon_process(ev, sc.tof_res_ps).items()}
    row_ptr, col, val = (t.cpu().numpy() for t in psf._fwd)
    xy = sparse.csr_matrix((val.astype(np.float64), col, row_ptr), shape=(1600, 1600))
    z = psf._mb.cpu().numpy().astype(np.float64)
    ref = OR.osem(ref_groups, np.ones(grid, np.float32), E.astype(np.float64), center, size, kw,
                  True, tb, ts2, 2, 3, psf=(xy, z))
    e = H.rel_l2(x.cpu().numpy(), ref)
    print(f"OSEM with PSF, 2 it x 3 subsets: rel_l2 vs f64 {e:.2e}")
    assert e <= 1e-4
