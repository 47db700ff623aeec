"""The TOR kernels at the large-frame pitch: 1.6 x 1.6 x 1.7 mm voxels,
kernel_width 4 mm, TOF 400 ps FWHM (sigma 25.48 mm).  There the reference's
patch is P = ceil((sqrt(2) 4 + 1.7) / 1.6) = 5 (tor.cu.cc:375) and the
kernels run their 4 x 4 candidate box (box 3 elsewhere): slice_box<..,4,..>,
the frame gathers of record_proj_kernel<4,..>, tile_scatter<..,4,..> and the
untiled tor_kernel<..,4,..>.  Every op is checked against the FP64 oracle
(rel L2 <= 1e-5, counts bit-exact), in the op frame and through the
reconstruction path of the world frame (OSEM 2 iterations x 2 subsets,
rel L2 <= 1e-4), including events whose box runs past the end of the image.
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL64 = 1e-5
GRID = (40, 36, 24)
CENTER = (1.0, -2.0, 0.5)
SIZE = (64.0, 57.6, 40.8)  # 1.6 x 1.6 x 1.7 mm
TOF_BIN = 6.0
TOF_SIGMA2 = (400.0 * 0.15 / 2.35482005) ** 2  # 649.2 mm^2
GEO = (GRID, CENTER, SIZE)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.fixture(scope="module")
def case(cuda):
    from srf_amd import _lib
    assert _lib.tor_patch_info(GRID, SIZE, H.KW) == (5, 4)
    lors = H.tor_lors(5000, seed=51, grid=GRID, center=CENTER, size=SIZE, full=True)
    img = H.image(GRID, seed=52)
    return lors, img


def _args(geo, tof):
    return (*GEO, H.KW, geo, TOF_BIN, TOF_SIGMA2 if tof else 1.0e6)


@pytest.mark.parametrize("geo,tof", [(False, True), (True, True), (True, False)])
@pytest.mark.parametrize("schedule", [None, (4, 256), (16, 128)])
def test_untiled_ops(cuda, case, geo, tof, schedule):
    import torch
    from oracle import lib as O
    from srf_amd import ops
    lors, img = case
    args = _args(geo, tof)
    L, X = _t(lors, cuda), _t(img, cuda)
    p = ops.projection(L, X, *args, schedule=schedule).cpu().numpy()
    ref = O.tor_project(lors, img, *args, prec="f64")
    e = H.rel_l2(p, ref)
    print(f"box 4 geo={geo} tof={tof} schedule={schedule}: proj rel_l2 {e:.2e}")
    assert e <= TOL64
    v = np.random.default_rng(3).uniform(0.5, 2.0, lors.shape[1]).astype(np.float32)
    v[::13] = 0.0
    v[1::31] = 1e-8
    b = ops.backprojection(X, L, _t(v, cuda), *GEO, args[5], args[6], H.KW, geo).cpu().numpy()
    rb = O.tor_backproject(lors, v, *args, prec="f64")
    assert H.rel_l2(b, rb) <= TOL64
    g = (v - 1.0).astype(np.float32)
    a = ops.backprojection_adjoint(X, L, _t(g, cuda), *GEO, args[5], args[6], H.KW,
                                   geo).cpu().numpy()
    ra = O.tor_backproject_adjoint(lors, g, *args, prec="f64")
    assert H.rel_l2(a, ra) <= TOL64
    f = ops.osem_subset(L, X, *args, schedule=schedule).cpu().numpy()
    keep = ref > 1e-7
    inv = np.where(keep, 1.0 / np.where(keep, ref, 1.0), 0.0).astype(np.float32)
    rf = O.tor_backproject_adjoint(lors, inv, *args)
    assert H.rel_l2(f, rf) <= TOL64
    # the reference's full P x P patch loop gives the box walk's bits
    pf = ops.projection(L, X, *args, full_patch=True).cpu().numpy()
    if schedule is None:
        assert np.array_equal(pf, p)
    torch.cuda.synchronize()


@pytest.mark.parametrize("geo", [False, True])
def test_count_and_map(cuda, case, geo):
    from oracle import lib as O
    from srf_amd import ops
    lors, img = case
    per, tot = ops.count_nnz(_t(lors, cuda), *GEO, H.KW, geo)
    ref = O.tor_count(lors, *GEO, H.KW, geo, prec="f64")
    assert np.array_equal(per.cpu().numpy(), ref)
    assert int(tot.item()) == int(ref.sum())
    print(f"box 4 count geo={geo}: {ref.sum() / len(ref):.1f} per event")
    l7 = np.ascontiguousarray(np.vstack([lors[0:6], lors[9:10]]))
    w = np.random.default_rng(5).uniform(0.5, 2.0, l7.shape[1]).astype(np.float32)
    X = _t(img, cuda)
    m = ops.maplors(X, *GEO, _t(l7, cuda), _t(w, cuda), H.KW, geo).cpu().numpy()
    r = O.tor_map(l7, w, *GEO, H.KW, geo, prec="f64")
    assert H.rel_l2(m, r) <= TOL64


@pytest.mark.parametrize("tile,chunk", [((8, 32, 20), 2048), ((4, 16, 16), 512),
                                        ((1, 5, 5), 64), ((16, 8, 8), 300)])
def test_plans(cuda, case, tile, chunk):
    """Tile plans at box 4: the tiled gather, the tiled scatters and the plan
    OSEM step (record projection from the contiguous op frame through buffer
    loads, then the tiled scatter of 1/p), and the scales it stores."""
    from oracle import lib as O
    from srf_amd.plan import TorPlan
    lors, img = case
    args = _args(True, True)
    L, X = _t(lors, cuda), _t(img, cuda)
    plan = TorPlan(L, *GEO, H.KW, True, TOF_BIN, TOF_SIGMA2, tile=tile, chunk=chunk,
                   tiled_projection=True)
    ref = O.tor_project(lors, img, *args, prec="f64")
    p = plan.projection(X).cpu().numpy()
    assert H.rel_l2(p, ref) <= TOL64
    v = np.random.default_rng(2).uniform(0.5, 2.0, lors.shape[1]).astype(np.float32)
    v[::11] = 0.0
    b = plan.backprojection(_t(v, cuda)).cpu().numpy()
    assert H.rel_l2(b, O.tor_backproject(lors, v, *args, prec="f64")) <= TOL64
    a = plan.backprojection_adjoint(_t(v - 1.0, cuda)).cpu().numpy()
    assert H.rel_l2(a, O.tor_backproject_adjoint(lors, v - 1.0, *args)) <= TOL64
    plan.close()
    plan = TorPlan(L, *GEO, H.KW, True, TOF_BIN, TOF_SIGMA2, tile=tile, chunk=chunk)
    f = plan.osem_subset(X).cpu().numpy()
    keep = ref > 1e-7
    inv = np.where(keep, 1.0 / np.where(keep, ref, 1.0), 0.0)
    rf = O.tor_backproject_adjoint(lors, inv.astype(np.float32), *args)
    e = H.rel_l2(f, rf)
    s = plan.scales().cpu().numpy()
    es = H.rel_l2(s, inv)
    print(f"box 4 plan {tile}/{chunk}: OSEM step rel_l2 {e:.2e}, scales 1/p rel_l2 {es:.2e}")
    assert e <= TOL64
    assert es <= TOL64
    assert np.array_equal(s == 0, ~keep)
    plan.close()


def test_single_walk_plans(cuda, case):
    """Events whose ellipse spans at most 3 rows (nearly parallel to the slice
    axis) all walk the 4 x 3 box: the plan's scatter then runs the kernel
    holding only that walk (tile 12 x 32 x 10: four workgroups per CU fit its
    LDS), and the general kernel where the tile is too large for that
    (12 x 32 x 16).  Both against the FP64 oracle."""
    from oracle import lib as O
    from srf_amd.plan import TorPlan
    lors, img = case
    lors = lors.copy()
    rng = np.random.default_rng(53)
    n = lors.shape[1]
    lors[3:5] = lors[0:2] + rng.uniform(-4.0, 4.0, (2, n)).astype(np.float32)
    r = rng.uniform(0.2, 0.8, n).astype(np.float32)
    lors[6:9] = lors[0:3] + r * (lors[3:6] - lors[0:3])
    # the geometric sigma factor scales the tube's radius: at most 0.6 keeps
    # every ellipse within 3 rows of 1.6 mm
    lors[9] = rng.uniform(0.3, 0.6, n).astype(np.float32)
    args = _args(True, True)
    L, X = _t(lors, cuda), _t(img, cuda)
    ref = O.tor_project(lors, img, *args, prec="f64")
    keep = ref > 1e-7
    inv = np.where(keep, 1.0 / np.where(keep, ref, 1.0), 0.0)
    rf = O.tor_backproject_adjoint(lors, inv.astype(np.float32), *args)
    v = np.random.default_rng(54).uniform(0.5, 2.0, n).astype(np.float32)
    rb = O.tor_backproject(lors, v, *args, prec="f64")
    for tile in ((12, 32, 10), (12, 32, 16)):
        plan = TorPlan(L, *GEO, H.KW, True, TOF_BIN, TOF_SIGMA2, tile=tile, chunk=1024)
        assert plan.walks() == (4, True)
        f = plan.osem_subset(X).cpu().numpy()
        b = plan.backprojection(_t(v, cuda)).cpu().numpy()
        plan.close()
        print(f"box 4 single-walk plan {tile}: OSEM step rel_l2 {H.rel_l2(f, rf):.2e}, "
              f"backprojection {H.rel_l2(b, rb):.2e}")
        assert H.rel_l2(f, rf) <= TOL64
        assert H.rel_l2(b, rb) <= TOL64


def test_plan_maplors(cuda, case):
    from oracle import lib as O
    from srf_amd.plan import TorPlan
    lors, _ = case
    l7 = np.ascontiguousarray(np.vstack([lors[0:6], lors[9:10]]))
    w = np.random.default_rng(7).uniform(0.5, 2.0, l7.shape[1]).astype(np.float32)
    plan = TorPlan(_t(l7, cuda), *GEO, H.KW, True, ncols=7, tile=(8, 16, 16), chunk=256)
    m = plan.maplors(_t(w, cuda)).cpu().numpy()
    assert H.rel_l2(m, O.tor_map(l7, w, *GEO, H.KW, True, prec="f64")) <= TOL64
    plan.close()


def test_far_corner(cuda):
    """Events through the last voxel rows and columns of the frame: the 4 x 4
    candidate boxes run past the image end (buffer loads read 0 there)."""
    from oracle import lib as O
    from srf_amd import ops
    from srf_amd.plan import TorPlan
    rng = np.random.default_rng(61)
    c, s = np.array(CENTER, np.float64), np.array(SIZE, np.float64)
    lo, hi = c - s / 2, c + s / 2
    n = 600
    p1 = np.empty((n, 3))
    p1[:, 0] = rng.uniform(hi[0] - 2.5, hi[0], n)
    p1[:, 1] = rng.uniform(hi[1] - 2.5, hi[1], n)
    p1[:, 2] = lo[2] - 5.0
    p2 = p1.copy()
    p2[:, 0] += rng.uniform(-1.5, 1.5, n)
    p2[:, 1] += rng.uniform(-1.5, 1.5, n)
    p2[:, 2] = hi[2] + 5.0
    pc = p1 + rng.uniform(0.75, 0.97, n)[:, None] * (p2 - p1)
    f = rng.uniform(0.5, 1.5, n)
    lors = np.vstack([p1.T, p2.T, pc.T, f[None, :]]).astype(np.float32)
    img = H.image(GRID, seed=62)
    args = _args(False, True)
    L, X = _t(lors, cuda), _t(img, cuda)
    plan = TorPlan(L, *GEO, H.KW, False, TOF_BIN, TOF_SIGMA2, tile=(8, 16, 16), chunk=256)
    got = plan.osem_subset(X).cpu().numpy()
    plan.close()
    ref_p = O.tor_project(lors, img, *args, prec="f64")
    keep = ref_p > 1e-7
    inv = np.where(keep, 1.0 / np.where(keep, ref_p, 1.0), 0.0).astype(np.float32)
    ref = O.tor_backproject_adjoint(lors, inv, *args)
    assert float(np.abs(ref[-1, -1, :]).sum()) > 0
    assert H.rel_l2(got, ref) <= TOL64
    assert H.rel_l2(ops.projection(L, X, *args).cpu().numpy(), ref_p) <= TOL64


def test_world_osem_vs_oracle(cuda):
    """Scanner events (400 ps TOF) through a 64 x 64 x 40 world image of
    1.6 x 1.6 x 1.7 mm: preprocessing, the x / y / z group frames, tile
    plans with op-frame gathers, 2 iterations x 2 subsets, against the
    oracle's OSEM (recon_step.py:33-49 over worker.py:79-85 subsets)."""
    import torch
    from oracle import recon as R
    from srf_amd import physics, preprocess, synth
    from srf_amd.data import Image, split_from_groups
    from srf_amd.recon import OSEM
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner(tof_res_ps=400.0)
    grid, center, size = (64, 64, 40), (0.0, 0.0, 0.0), (102.4, 102.4, 68.0)
    ph = synth.Phantom(45.0, 30.0, [((10.0, -5.0, 3.0), 12.0, 4.0)])
    ev, _ = synth.tof_events(30000, ph, seed=63, scanner=sc, device=cuda)
    groups = preprocess.recon_process(ev, sc.tof_res_ps)
    tb, ts2 = sc.tof_info()
    for a in "xy":
        assert groups[a].shape[1] > 0
        from srf_amd._lib import tor_patch_info
        g, c, s, _ = physics.group_frame(grid, center, size, (1, 1, 1), a)
        assert tor_patch_info(g, s, 4.0)[1] == 4
    model = physics.SplitLoRsModel(4.0, False, tof_sigma2=ts2, tof_bin=tb)
    model.schedule["projection"] = (4, 256)
    rng = np.random.default_rng(64)
    emap = rng.uniform(0.5, 1.5, grid).astype(np.float32)
    x0 = np.ones(grid, np.float32)
    x = torch.from_numpy(x0).to(cuda)
    osem = OSEM(model, split_from_groups(groups), Image(torch.from_numpy(emap).to(cuda), center,
                                                        size), nb_subsets=2, tiled=True)
    osem.iterate(Image(x, center, size), 2)
    got = x.cpu().numpy()
    osem.close()
    gnp = {a: groups[a].cpu().numpy() for a in "xyz"}
    ref = R.osem(gnp, x0, emap.astype(np.float64), center, size, 4.0, False, tb, ts2, 2, 2)
    e = H.rel_l2(got, ref)
    print(f"box 4 world OSEM 2x2: {sum(gnp[a].shape[1] for a in 'xyz')} events, rel_l2 {e:.2e}")
    assert e <= 1e-4
