"""GPU parity of the tiled Siddon backprojection (srfx_siddon_plan_*).

A plan resumes each LOR's walk per tile from the state the full walk reached,
so every run visits the voxels and forms the segment values of the untiled
walk; only the summation order changes (FP64 per tile, then one FP32 atomic
per voxel and chunk).  Tolerances: rel L2 <= 1e-6 against the untiled HIP
ops and <= 1e-5 against the FP64 oracle (siddon.cu.cc:339-381 semantics);
OSEM through the driver <= 1e-4 against the oracle's OSEM."""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

GEO = (H.SID_GRID, H.SID_CENTER, H.SID_SIZE)
# ops-level images are [gz][gy][gx] arrays: x is the fastest axis
GX, GY, GZ = H.SID_GRID
OPS_STRIDES = (1, GX, GX * GY)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.fixture(scope="module")
def case(cuda):
    return H.siddon_lors(4000, seed=3), H.image(H.SID_GRID, seed=3)


def _plan(lors, dev, ts2, tile, chunk=2048, block=256):
    from srf_amd.plan import SiddonPlan
    return SiddonPlan(_t(lors, dev), *GEO, 6.0, ts2, strides=OPS_STRIDES, tile=tile,
                      chunk=chunk, block=block)


# tiles that divide the grid, that do not, a single tile, one-voxel tiles
TILES = [((8, 8, 8), 2048, 256), ((7, 5, 16), 100, 64), ((24, 20, 16), 2048, 256),
         ((1, 1, 1), 512, 128), ((16, 16, 16), 2048, 256)]


@pytest.mark.parametrize("tof", [True, False])
@pytest.mark.parametrize("tile,chunk,block", TILES)
def test_backproject(cuda, case, tof, tile, chunk, block):
    from oracle import lib as O
    from srf_amd import ops
    lors, img = case
    ts2 = 400.0 if tof else 1.0e6
    rng = np.random.default_rng(2)
    v = rng.uniform(0.5, 2.0, lors.shape[1]).astype(np.float32)
    v[::13] = 0.0
    v[5::17] = -1.0
    plan = _plan(lors, cuda, ts2, tile, chunk, block)
    nr, nc = plan.info()
    out = _t(np.zeros_like(img), cuda)
    plan.backprojection(_t(v, cuda), out=out)
    b = out.cpu().numpy()
    u = ops.siddon_backprojection(_t(img, cuda), _t(lors, cuda), _t(v, cuda), *GEO, 6.0,
                                  ts2).cpu().numpy()
    r64 = O.siddon_backproject(lors, v, *GEO, 6.0, ts2, prec="f64")
    eu, e64 = H.rel_l2(b, u), H.rel_l2(b, r64)
    print(f"siddon plan bp tof={tof} tile={tile}: {nr} runs {nc} chunks, vs untiled {eu:.2e} "
          f"vs f64 {e64:.2e}")
    assert nr > 0 and nc > 0
    assert eu <= 1e-6 and e64 <= 1e-5
    plan.close()


def test_accumulate_and_adjoint(cuda, case):
    import torch
    from srf_amd import ops
    lors, img = case
    plan = _plan(lors, cuda, 400.0, (8, 8, 8))
    y = np.random.default_rng(4).normal(0, 1, lors.shape[1]).astype(np.float32)
    y[::7] = 0.0
    L, X, Y = _t(lors, cuda), _t(img, cuda), _t(y, cuda)
    base = _t(np.full_like(img, 0.25), cuda)
    out = base.clone()
    plan.backprojection_adjoint(Y, out=out, accumulate=True)
    u = ops.siddon_backprojection_adjoint(X, L, Y, *GEO, 6.0, 400.0)
    e_acc = H.rel_l2(out.cpu().numpy(), (base + u).cpu().numpy())
    ATy = plan.backprojection_adjoint(Y, out=torch.empty_like(base))
    e_adj = H.rel_l2(ATy.cpu().numpy(), u.cpu().numpy())
    Ax = ops.siddon_projection(L, X, *GEO, 6.0, 400.0)
    lhs = float(torch.dot(Ax.double(), Y.double()))
    rhs = float(torch.dot(X.reshape(-1).double(), ATy.reshape(-1).double()))
    e_ip = abs(lhs - rhs) / abs(lhs)
    print(f"siddon plan adjoint: accumulate {e_acc:.2e}, vs untiled {e_adj:.2e}, "
          f"<Ax,y> vs <x,A^T y> {e_ip:.2e}")
    assert e_acc <= 1e-6 and e_adj <= 1e-6 and e_ip <= 1e-5
    plan.close()


@pytest.mark.parametrize("tof", [True, False])
def test_osem_subset(cuda, case, tof):
    from oracle import lib as O
    from srf_amd import ops
    lors, img = case
    ts2 = 400.0 if tof else 1.0e6
    plan = _plan(lors, cuda, ts2, (8, 8, 8))
    X = _t(img, cuda)
    out = _t(np.zeros_like(img), cuda)
    plan.osem_subset(X, out=out)
    f = ops.siddon_osem_subset(_t(lors, cuda), X, *GEO, 6.0, ts2)
    p = plan.projection().cpu().numpy()
    p64 = O.siddon_project(lors, img, *GEO, 6.0, ts2, prec="f64")
    r64 = O.siddon_backproject(lors, p64.astype(np.float32), *GEO, 6.0, ts2, prec="f64")
    ef, ep, e64 = (H.rel_l2(out.cpu().numpy(), f.cpu().numpy()), H.rel_l2(p, p64),
                   H.rel_l2(out.cpu().numpy(), r64))
    print(f"siddon plan osem tof={tof}: vs fused {ef:.2e}, projection vs f64 {ep:.2e}, "
          f"bp(1/p) vs f64 {e64:.2e}")
    assert ef <= 1e-6 and ep <= 1e-5 and e64 <= 1e-5
    plan.close()


def test_world_layout_and_empty(cuda):
    """A (nx, ny, nz) world image (z fastest), as the Siddon model passes it,
    and plans of no LORs / of LORs that miss the image."""
    import torch
    from srf_amd import physics
    from srf_amd.data import Image, ListModeData
    from srf_amd.plan import SiddonPlan
    grid, center, size = (20, 18, 12), (1.0, -2.0, 0.5), (60.0, 54.0, 40.8)
    lors = H.siddon_lors(3000, seed=9, grid=grid, center=center, size=size)
    rng = np.random.default_rng(10)
    v = rng.uniform(0.5, 2.0, lors.shape[1]).astype(np.float32)
    model = physics.CompleteLoRsModel("siddon", tof_sigma2=1.0e6, tof_bin=6.0)
    x = torch.zeros(grid, device=cuda)
    data = ListModeData(_t(lors, cuda), _t(v, cuda), soa=True)
    plans = model.make_plans(Image(x, center, size), data)
    b = model.backprojection(data, Image(x, center, size), plans=plans).data
    u = model.backprojection(data, Image(x, center, size)).data
    assert H.rel_l2(b.cpu().numpy(), u.cpu().numpy()) <= 1e-6
    plans["siddon"].close()
    empty = SiddonPlan(torch.zeros(7, 0, device=cuda), grid, center, size, 6.0, 1.0e6)
    assert empty.info() == (0, 0)
    assert float(empty.backprojection(torch.zeros(0, device=cuda)).abs().sum()) == 0.0
    empty.close()
    far = np.zeros((7, 64), np.float32)
    far[0], far[3] = 500.0, 600.0          # x from 500 to 600 mm: outside the box
    far[1], far[4] = np.linspace(-20, 20, 64), np.linspace(-20, 20, 64)
    miss = SiddonPlan(_t(far, cuda), grid, center, size, 6.0, 1.0e6)
    assert miss.info() == (0, 0)
    miss.close()


@pytest.mark.parametrize("fused", [False, True])
def test_osem_driver_vs_oracle(cuda, fused):
    """Non-TOF Siddon OSEM, 2 iterations x 4 subsets, through the driver with
    tiled plans against the oracle's OSEM: rel L2 <= 1e-4."""
    import torch
    from oracle import recon as R
    from srf_amd import physics
    from srf_amd.data import Image, ListModeData
    from srf_amd.recon import OSEM
    grid, center, size = (32, 32, 16), (0.5, -1.0, 0.25), (96.0, 96.0, 54.4)
    lors = H.siddon_lors(12000, seed=17, grid=grid, center=center, size=size)
    rng = np.random.default_rng(18)
    emap = rng.uniform(0.5, 1.5, grid).astype(np.float32)
    x0 = np.ones(grid, np.float32)
    model = physics.CompleteLoRsModel("siddon", tof_sigma2=1.0e6, tof_bin=6.0)
    data = ListModeData(_t(lors, cuda), soa=True)
    x = _t(x0, cuda)
    osem = OSEM(model, data, Image(_t(emap, cuda), center, size), nb_subsets=4, fused=fused,
                tiled=True)
    osem.iterate(Image(x, center, size), 2)
    torch.cuda.synchronize()
    assert osem.plans is not None and len(osem.plans) == 4
    osem.close()
    ref = R.siddon_osem(lors, x0, emap.astype(np.float64), center, size, 6.0, 1.0e6, 2, 4)
    e = H.rel_l2(x.cpu().numpy(), ref)
    print(f"tiled Siddon OSEM 2 it x 4 subsets fused={fused}: rel_l2 {e:.2e}")
    assert e <= 1e-4


def test_subset_plans_equal_single_plans(cuda, case):
    """srfx_siddon_plan_create_subsets builds, for each LOR range, the plan
    srfx_siddon_plan_create builds for it: same runs and chunks, and the same
    backprojection; empty ranges included."""
    from srf_amd.plan import SiddonPlan
    lors, img = case
    n = lors.shape[1]
    bounds = [0, 700, 700, 2100, 3999, n]
    L = _t(lors, cuda)
    rng = np.random.default_rng(7)
    v = _t(rng.uniform(0.5, 2.0, n).astype(np.float32), cuda)
    set_plans = SiddonPlan.subsets(L, bounds, *GEO, 6.0, 400.0, strides=OPS_STRIDES,
                                   tile=(8, 8, 8), chunk=300)
    for k, pl in enumerate(set_plans):
        lo, hi = bounds[k], bounds[k + 1]
        one = SiddonPlan(L[:, lo:hi], *GEO, 6.0, 400.0, strides=OPS_STRIDES, tile=(8, 8, 8),
                         chunk=300)
        assert pl.info() == one.info()
        a = pl.backprojection(v[lo:hi], out=_t(np.zeros_like(img), cuda)).cpu().numpy()
        b = one.backprojection(v[lo:hi], out=_t(np.zeros_like(img), cuda)).cpu().numpy()
        if hi > lo:
            assert H.rel_l2(a, b) <= 1e-6
        else:
            assert not a.any() and not b.any()
        one.close()
    # plans of a set are freed one by one, in any order
    for pl in reversed(set_plans):
        pl.close()


@pytest.mark.parametrize("tof", [True, False])
def test_sorted_stream(cuda, case, tof):
    """The sorted LOR stream changes only the order of the sums: the same
    runs, backprojections equal to the unsorted plan's, and the OSEM step's
    projection, returned in the caller's LOR order, equal bit for bit to the
    untiled projection (each LOR is traced from the same bits)."""
    from srf_amd import ops
    from srf_amd.plan import SiddonPlan
    lors, img = case
    ts2 = 400.0 if tof else 1.0e6
    L, X = _t(lors, cuda), _t(img, cuda)
    rng = np.random.default_rng(11)
    v = _t(rng.uniform(0.5, 2.0, lors.shape[1]).astype(np.float32), cuda)
    a = SiddonPlan(L, *GEO, 6.0, ts2, strides=OPS_STRIDES, tile=(8, 8, 8), sort=True)
    b = SiddonPlan(L, *GEO, 6.0, ts2, strides=OPS_STRIDES, tile=(8, 8, 8), sort=False)
    assert a.info()[0] == b.info()[0]
    ba = a.backprojection(v, out=_t(np.zeros_like(img), cuda)).cpu().numpy()
    bb = b.backprojection(v, out=_t(np.zeros_like(img), cuda)).cpu().numpy()
    assert H.rel_l2(ba, bb) <= 1e-6
    a.osem_subset(X, out=_t(np.zeros_like(img), cuda))
    p = ops.siddon_projection(L, X, *GEO, 6.0, ts2)
    assert np.array_equal(a.projection().cpu().numpy(), p.cpu().numpy())
    a.close()
    b.close()


def test_osem_counts_and_stages(cuda, case):
    """The OSEM step with per-LOR counts backprojects c/p, as the untiled
    backprojection of lors_value = p / c does (list-mode data with counts);
    the step run as its two halves equals the one call."""
    from srf_amd import ops
    lors, img = case
    L, X = _t(lors, cuda), _t(img, cuda)
    rng = np.random.default_rng(12)
    c = rng.poisson(3.0, lors.shape[1]).astype(np.float32)
    C = _t(c, cuda)
    plan = _plan(lors, cuda, 1.0e6, (8, 8, 8))
    full = plan.osem_subset(X, out=_t(np.zeros_like(img), cuda), counts=C).cpu().numpy()
    plan.osem_subset(X, stage="project")
    two = plan.osem_subset(None, out=_t(np.zeros_like(img), cuda), counts=C,
                           stage="scatter").cpu().numpy()
    p = ops.siddon_projection(L, X, *GEO, 6.0, 1.0e6)
    ref = ops.siddon_backprojection(X, L, p / C, *GEO, 6.0, 1.0e6).cpu().numpy()
    e_ref, e_two = H.rel_l2(full, ref), H.rel_l2(two, full)
    print(f"siddon plan osem with counts: vs untiled bp(p / c) {e_ref:.2e}, two halves {e_two:.2e}")
    assert e_ref <= 1e-6 and e_two <= 1e-6
    plan.close()
