"""GPU parity of the Siddon ops against the CPU oracle (FP64 restatement of
siddon.cu.cc and the FP32 port).  The kernels trace rays in FP64, so voxel
decisions (including rays lying exactly on voxel faces) are those of the FP64
oracle.  Tolerances: rel L2 <= 1e-5 vs FP64 (measured ~1e-7); segment counts
bit-exact vs FP64; vs the FP32 port rel L2 <= 1e-5 on rays off voxel faces
(on a face the reference's FP32 rounding picks the voxel)."""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.fixture(scope="module")
def case(cuda):
    return H.siddon_lors(4000, seed=3), H.image(H.SID_GRID, seed=3)


GEO = (H.SID_GRID, H.SID_CENTER, H.SID_SIZE)


@pytest.mark.parametrize("tof", [True, False])
def test_project(cuda, case, tof):
    from oracle import lib as O
    from srf_amd import ops
    lors, img = case
    ts2 = 400.0 if tof else 1.0e6
    p = ops.siddon_projection(_t(lors, cuda), _t(img, cuda), *GEO, 6.0, ts2).cpu().numpy()
    r64 = O.siddon_project(lors, img, *GEO, 6.0, ts2, prec="f64")
    e64 = H.rel_l2(p, r64)
    gen = H.siddon_lors(4000, seed=3, special=False)
    pg = ops.siddon_projection(_t(gen, cuda), _t(img, cuda), *GEO, 6.0, ts2).cpu().numpy()
    e32 = H.rel_l2(pg, O.siddon_project(gen, img, *GEO, 6.0, ts2, prec="f32"))
    print(f"siddon project tof={tof}: vs f64 {e64:.2e} vs f32 (off-face rays) {e32:.2e}")
    assert e64 <= 1e-5 and e32 <= 1e-5


@pytest.mark.parametrize("tof", [True, False])
def test_backproject(cuda, case, tof):
    from oracle import lib as O
    from srf_amd import ops
    lors, img = case
    ts2 = 400.0 if tof else 1.0e6
    rng = np.random.default_rng(2)
    v = rng.uniform(0.5, 2.0, lors.shape[1]).astype(np.float32)
    v[::13] = 0.0
    b = ops.siddon_backprojection(_t(img, cuda), _t(lors, cuda), _t(v, cuda), *GEO, 6.0,
                                  ts2).cpu().numpy()
    r64 = O.siddon_backproject(lors, v, *GEO, 6.0, ts2, prec="f64")
    e64 = H.rel_l2(b, r64)
    print(f"siddon bp tof={tof}: vs f64 {e64:.2e}")
    assert e64 <= 1e-5


def test_map(cuda, case):
    from oracle import lib as O
    from srf_amd import ops
    lors, img = case
    v = np.ones(lors.shape[1], np.float32)
    m = ops.siddon_maplors(_t(img, cuda), *GEO, _t(lors, cuda), _t(v, cuda)).cpu().numpy()
    r64 = O.siddon_map(lors, v, *GEO, prec="f64")
    assert H.rel_l2(m, r64) <= 1e-5


@pytest.mark.parametrize("tof", [True, False])
def test_count_exact(cuda, case, tof):
    from oracle import lib as O
    from srf_amd import ops
    lors, _ = case
    ts2 = 400.0 if tof else 1.0e6
    per, tot = ops.siddon_count_nnz(_t(lors, cuda), *GEO, 6.0, ts2)
    ref = O.siddon_count(lors, *GEO, 6.0, ts2, prec="f64")
    assert np.array_equal(per.cpu().numpy(), ref)
    assert int(tot.item()) == int(ref.sum())


def test_adjoint(cuda, case):
    import torch
    from srf_amd import ops
    lors, img = case
    y = np.random.default_rng(4).normal(0, 1, lors.shape[1]).astype(np.float32)
    L, X, Y = _t(lors, cuda), _t(img, cuda), _t(y, cuda)
    Ax = ops.siddon_projection(L, X, *GEO, 6.0, 400.0)
    ATy = ops.siddon_backprojection_adjoint(X, L, Y, *GEO, 6.0, 400.0)
    lhs = float(torch.dot(Ax.double(), Y.double()))
    rhs = float(torch.dot(X.reshape(-1).double(), ATy.reshape(-1).double()))
    assert abs(lhs - rhs) / abs(lhs) <= 1e-5


def test_osem_fused(cuda, case):
    from srf_amd import ops
    lors, img = case
    L, X = _t(lors, cuda), _t(img, cuda)
    p = ops.siddon_projection(L, X, *GEO, 6.0, 1.0e6)
    b = ops.siddon_backprojection(X, L, p, *GEO, 6.0, 1.0e6)
    f = ops.siddon_osem_subset(L, X, *GEO, 6.0, 1.0e6)
    assert H.rel_l2(f.cpu().numpy(), b.cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("fused", [False, True])
def test_osem_vs_oracle(cuda, fused):
    """Non-TOF Siddon OSEM, 2 iterations x 4 subsets, through the
    reconstruction driver on the Siddon model (complete_lors.py:59-98 +
    recon_step.py:33-49: world image, [2, 1, 0] frame, contiguous subsets)
    against the oracle's OSEM: rel L2 <= 1e-4."""
    import torch
    from oracle import recon as R
    from srf_amd import physics
    from srf_amd.data import Image, ListModeData
    from srf_amd.recon import OSEM
    grid, center, size = (32, 32, 16), (0.5, -1.0, 0.25), (96.0, 96.0, 54.4)  # 3x3x3.4 mm
    lors = H.siddon_lors(12000, seed=17, grid=grid, center=center, size=size)
    rng = np.random.default_rng(18)
    emap = rng.uniform(0.5, 1.5, grid).astype(np.float32)
    x0 = np.ones(grid, np.float32)
    model = physics.CompleteLoRsModel("siddon", tof_sigma2=1.0e6, tof_bin=6.0)
    data = ListModeData(_t(lors, cuda), soa=True)
    x = _t(x0, cuda)
    osem = OSEM(model, data, Image(_t(emap, cuda), center, size), nb_subsets=4, fused=fused)
    osem.iterate(Image(x, center, size), 2)
    torch.cuda.synchronize()
    ref = R.siddon_osem(lors, x0, emap.astype(np.float64), center, size, 6.0, 1.0e6, 2, 4)
    e = H.rel_l2(x.cpu().numpy(), ref)
    print(f"Siddon OSEM 2 it x 4 subsets fused={fused}: {lors.shape[1]} LORs, rel_l2 {e:.2e}")
    assert e <= 1e-4
