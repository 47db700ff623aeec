"""Tiled TOR ops (srfx_tor_plan_*) against the FP64 oracle.

A plan bins each event into the image tiles it can touch; the tiled
scatter accumulates per tile in LDS and the tiled gather reads tiles staged in
LDS.  Every op must give the untiled results: rel L2 <= 1e-5 against the FP64
oracle for every tile shape, including windows narrower than the patch (an
event then spans many tiles) and chunk sizes that split a tile over many
workgroups.  The tiled projection sums each event's per-tile partials in a
fixed order, so it is deterministic.
"""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL64 = 1e-5
# (slices, wx, wy), chunk.  (1, 2, 2) has more tiles than the per-workgroup
# LDS counters of the plan build hold (global-atomic build path).
TILES = [((8, 32, 32), 4096), ((1, 4, 4), 64), ((3, 5, 7), 97), ((16, 16, 32), 1024),
         ((1, 2, 2), 50)]


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.fixture(scope="module")
def case(cuda):
    lors = H.tor_lors(6000, seed=11)
    img = H.image(H.TOR_GRID, seed=4)
    return lors, img


def _plan(L, geo, tof, tile, chunk, ncols=10, tiled_projection=True):
    from srf_amd.plan import TorPlan
    ts2 = H.TOF_SIGMA2 if tof else 1.0e6
    return TorPlan(L, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, geo, H.TOF_BIN, ts2,
                   ncols=ncols, tile=tile, chunk=chunk, tiled_projection=tiled_projection)


@pytest.mark.parametrize("tile,chunk", TILES)
@pytest.mark.parametrize("geo,tof", [(False, True), (True, True), (True, False)])
def test_plan_project_backproject(cuda, case, tile, chunk, geo, tof):
    from oracle import lib as O
    lors, img = case
    ts2 = H.TOF_SIGMA2 if tof else 1.0e6
    args = (H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, geo, H.TOF_BIN, ts2)
    L, X = _t(lors, cuda), _t(img, cuda)
    plan = _plan(L, geo, tof, tile, chunk)
    nl, nc = plan.info()
    assert nl >= lors.shape[1] // 2 and nc >= 1
    p = plan.projection(X).cpu().numpy()
    ref = O.tor_project(lors, img, *args, prec="f64")
    e = H.rel_l2(p, ref)
    print(f"plan {tile} chunk {chunk} geo={geo} tof={tof}: {nl} listings, {nc} chunks, "
          f"proj rel_l2 {e:.2e}")
    assert e <= TOL64
    assert np.array_equal(p, plan.projection(X).cpu().numpy())
    v = np.random.default_rng(2).uniform(0.5, 2.0, lors.shape[1]).astype(np.float32)
    v[::11] = 0.0
    v[1::29] = 1e-8
    b = plan.backprojection(_t(v, cuda)).cpu().numpy()
    rb = O.tor_backproject(lors, v, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, geo, H.TOF_BIN,
                           ts2, prec="f64")
    assert H.rel_l2(b, rb) <= TOL64
    a = plan.backprojection_adjoint(_t(v - 1.0, cuda)).cpu().numpy()
    ra = O.tor_backproject_adjoint(lors, v - 1.0, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW,
                                   geo, H.TOF_BIN, ts2)
    assert H.rel_l2(a, ra) <= TOL64
    plan.close()


# The projection from the event records (the OSEM step's walk), on the
# image in place and on the contiguous op-frame copy (frame gathers).
@pytest.mark.parametrize("geo,tof", [(True, True), (False, True), (True, False)])
def test_plan_record_projection(cuda, case, geo, tof):
    from oracle import lib as O
    lors, img = case
    ts2 = H.TOF_SIGMA2 if tof else 1.0e6
    args = (H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, geo, H.TOF_BIN, ts2)
    L, X = _t(lors, cuda), _t(img, cuda)
    assert X.is_contiguous() and tuple(X.shape) == tuple(H.TOR_GRID[::-1])
    ref = O.tor_project(lors, img, *args, prec="f64")
    for tiled in (False, True):
        plan = _plan(L, geo, tof, (8, 32, 32), 2048, tiled_projection=tiled)
        a = plan.projection(X, op_frame=True).cpu().numpy()
        assert H.rel_l2(a, ref) <= TOL64
        if not tiled:
            b = plan.projection(X).cpu().numpy()
            assert H.rel_l2(b, ref) <= TOL64
            assert np.abs(a - b).max() <= 1e-6 * np.abs(ref).max()
        plan.close()


# Plans without the tiled projection's slot tables sort their chunks and
# carry each listing's slices in its list entry when the slab has at most 16
# slices; a 20-slice slab sorts without them.
@pytest.mark.parametrize("tiled_projection,tile", [(False, (4, 16, 16)), (True, (4, 16, 16)),
                                                   (False, (20, 8, 8))])
def test_plan_osem_subset(cuda, case, tiled_projection, tile):
    import torch
    from oracle import lib as O
    from srf_amd import ops
    lors, img = case
    args = (H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, True, H.TOF_BIN, H.TOF_SIGMA2)
    L, X = _t(lors, cuda), _t(img, cuda)
    plan = _plan(L, True, True, tile, 512, tiled_projection=tiled_projection)
    f = plan.osem_subset(X, tiled_projection=tiled_projection).cpu().numpy()
    ref = O.tor_project(lors, img, *args, prec="f64")
    keep = ref > 1e-7
    inv = np.where(keep, 1.0 / np.where(keep, ref, 1.0), 0.0).astype(np.float32)
    rf = O.tor_backproject_adjoint(lors, inv, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, True,
                                   H.TOF_BIN, H.TOF_SIGMA2)
    assert H.rel_l2(f, rf) <= TOL64
    # accumulate adds into the given image
    base = torch.full_like(X, 0.5)
    acc = plan.osem_subset(X, out=base.clone(), accumulate=True,
                           tiled_projection=tiled_projection)
    un = ops.osem_subset(L, X, *args)
    assert H.rel_l2((acc - 0.5).cpu().numpy(), un.cpu().numpy()) <= 1e-6
    plan.close()


def test_plan_maplors(cuda, case):
    from oracle import lib as O
    lors, _ = case
    l7 = np.ascontiguousarray(np.vstack([lors[0:6], lors[9:10]]))
    w = np.random.default_rng(3).uniform(0.5, 2.0, l7.shape[1]).astype(np.float32)
    plan = _plan(_t(l7, cuda), True, False, (8, 16, 16), 256, ncols=7)
    m = plan.maplors(_t(w, cuda)).cpu().numpy()
    r = O.tor_map(l7, w, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, True, prec="f64")
    assert H.rel_l2(m, r) <= TOL64
    plan.close()


def test_single_walk_modes(cuda, case):
    """Events nearly parallel to the slice axis at kernel width 3.6 mm (a
    tube radius of 1.8 mm, under one 2 mm row) all walk the 3 x 2 box, so a
    plan of them (tile 12 x 32 x 10) scatters through the single-walk
    kernel in every mode: the backprojection, its adjoint, Maplors and the
    OSEM step, each against the FP64 oracle."""
    from oracle import lib as O
    from srf_amd.plan import TorPlan
    lors, img = case
    lors = lors.copy()
    rng = np.random.default_rng(61)
    n = lors.shape[1]
    lors[3:5] = lors[0:2] + rng.uniform(-4.0, 4.0, (2, n)).astype(np.float32)
    r = rng.uniform(0.2, 0.8, n).astype(np.float32)
    lors[6:9] = lors[0:3] + r * (lors[3:6] - lors[0:3])
    kw = 3.6
    geo = (H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE)
    args = (*geo, kw, False, H.TOF_BIN, H.TOF_SIGMA2)
    L, X = _t(lors, cuda), _t(img, cuda)
    plan = TorPlan(L, *geo, kw, False, H.TOF_BIN, H.TOF_SIGMA2, tile=(12, 32, 10), chunk=1024)
    assert plan.walks() == (3, True)
    v = rng.uniform(0.5, 2.0, n).astype(np.float32)
    b = plan.backprojection(_t(v, cuda)).cpu().numpy()
    assert H.rel_l2(b, O.tor_backproject(lors, v, *args, prec="f64")) <= TOL64
    a = plan.backprojection_adjoint(_t(v - 1.0, cuda)).cpu().numpy()
    assert H.rel_l2(a, O.tor_backproject_adjoint(lors, v - 1.0, *args)) <= TOL64
    f = plan.osem_subset(X).cpu().numpy()
    ref = O.tor_project(lors, img, *args, prec="f64")
    keep = ref > 1e-7
    inv = np.where(keep, 1.0 / np.where(keep, ref, 1.0), 0.0).astype(np.float32)
    assert H.rel_l2(f, O.tor_backproject_adjoint(lors, inv, *args)) <= TOL64
    plan.close()
    l7 = np.ascontiguousarray(np.vstack([lors[0:6], lors[9:10]]))
    plan = TorPlan(_t(l7, cuda), *geo, kw, False, ncols=7, tile=(12, 32, 10), chunk=1024)
    assert plan.walks() == (3, True)
    m = plan.maplors(_t(v, cuda)).cpu().numpy()
    assert H.rel_l2(m, O.tor_map(l7, v, *geo, kw, False, prec="f64")) <= TOL64
    plan.close()


def test_plan_empty(cuda):
    import torch
    from srf_amd.plan import TorPlan
    L = torch.zeros(10, 0, device=cuda)
    plan = TorPlan(L, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, tiled_projection=True)
    assert plan.info() == (0, 0)
    X = torch.ones(H.TOR_GRID[::-1], device=cuda)
    assert plan.projection(X).numel() == 0
    b = plan.backprojection(torch.zeros(0, device=cuda))
    assert float(b.abs().sum()) == 0.0
    plan.close()


def test_plan_without_projection_tables(cuda, case):
    """A plan built without the tiled projection's tables refuses the OSEM
    step's tiled projection, projects from its event records instead, and
    backprojects like the untiled op."""
    from srf_amd import ops
    from srf_amd._lib import SRFX_EUNSUPPORTED, SrfxError
    lors, img = case
    L, X = _t(lors, cuda), _t(img, cuda)
    plan = _plan(L, False, True, (8, 16, 16), 1024, tiled_projection=False)
    with pytest.raises(SrfxError) as ei:
        plan.osem_subset(X, tiled_projection=True)
    assert ei.value.status == SRFX_EUNSUPPORTED
    p = ops.projection(L, X, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, False, H.TOF_BIN,
                       H.TOF_SIGMA2)
    assert H.rel_l2(plan.projection(X).cpu().numpy(), p.cpu().numpy()) <= 1e-6
    b0 = ops.backprojection(X, L, p, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.TOF_BIN,
                            H.TOF_SIGMA2, H.KW, False)
    b1 = plan.backprojection(p)
    assert H.rel_l2(b1.cpu().numpy(), b0.cpu().numpy()) <= 1e-6
    plan.close()


def test_plan_osem_stages(cuda, case):
    """The OSEM step run as its two stages (projection into the plan's event
    records, then the tiled scatter) equals the single call."""
    import torch
    lors, img = case
    L, X = _t(lors, cuda), _t(img, cuda)
    from srf_amd import ops
    plan = _plan(L, False, True, (8, 16, 16), 512, tiled_projection=False)
    full = plan.osem_subset(X)
    ref = ops.osem_subset(L, X, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, False, H.TOF_BIN,
                          H.TOF_SIGMA2)
    assert float(ref.abs().sum()) > 0
    assert H.rel_l2(full.cpu().numpy(), ref.cpu().numpy()) <= 1e-6
    out = torch.full_like(X, 0.25)
    plan.osem_subset(X, out=out, accumulate=True, stage="project")
    assert float((out - 0.25).abs().max()) == 0.0  # the projection stage writes no image
    plan.osem_subset(X, out=out, accumulate=True, stage="scatter")
    assert H.rel_l2((out - 0.25).cpu().numpy(), full.cpu().numpy()) <= 1e-6
    from srf_amd._lib import SrfxError
    with pytest.raises(SrfxError):
        plan.osem_subset(X, stage="project", tiled_projection=True)
    plan.close()


@pytest.mark.parametrize("axis", ["x", "y"])
def test_plan_osem_op_frame_copy(cuda, case, axis):
    """A plan built for a strided (world-layout) image: the OSEM step reading
    the contiguous op-frame copy (SRFX_IMAGE_OP_FRAME) equals the step
    reading the world image in place, into the same world-layout output."""
    import torch
    from srf_amd import physics
    from srf_amd._lib import SrfxError
    from srf_amd.plan import TorPlan
    lors, img = case
    gx, gy, gz = H.TOR_GRID
    # a world image whose group-`axis` frame is the op frame of the case
    inv = {a: np.argsort(physics.PERM[a]) for a in "xy"}
    frame = torch.from_numpy(img).to(cuda)                 # [gz][gy][gx]
    world = frame.permute(*inv[axis]).contiguous()         # (nx, ny, nz)
    assert torch.equal(physics.SplitLoRsModel.frame_copy(world, axis), frame)
    g, c, s, st = physics.group_frame(world.shape, (0, 0, 0), (1, 1, 1), world.stride(), axis)
    assert g == [gx, gy, gz]
    L = _t(lors, cuda)
    plan = TorPlan(L, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, True, H.TOF_BIN, H.TOF_SIGMA2,
                   strides=st, tile=(8, 16, 16), chunk=512, axis=axis)
    a = plan.osem_subset(world, out=torch.zeros_like(world))
    b = plan.osem_subset(frame, out=torch.zeros_like(world), op_frame=True)
    assert float(a.abs().sum()) > 0
    assert H.rel_l2(b.cpu().numpy(), a.cpu().numpy()) <= 1e-6
    with pytest.raises(ValueError):
        plan.osem_subset(world, out=torch.zeros_like(world), op_frame=True)  # wrong shape
    with pytest.raises(SrfxError):
        plan.osem_subset(frame, out=torch.zeros_like(world), op_frame=True,
                         tiled_projection=True)
    plan.close()


def test_frame_gather_far_corner(cuda):
    """Events through the op frame's last voxel rows: the record projection
    reads a contiguous op frame with buffer loads whose candidate boxes run
    past the image end (those candidates read 0 and are dropped).  Voxels
    kept next to them must still read their values: the plan's OSEM step
    equals the untiled fused step (strided gathers), and both match the FP64
    oracle."""
    from oracle import lib as O
    from srf_amd import ops
    rng = np.random.default_rng(21)
    c = np.array(H.TOR_CENTER, np.float64)
    s = np.array(H.TOR_SIZE, np.float64)
    lo, hi = c - s / 2, c + s / 2
    n = 512
    p1 = np.empty((n, 3))
    p1[:, 0] = rng.uniform(hi[0] - 3.0, hi[0], n)   # last voxel columns along x
    p1[:, 1] = rng.uniform(hi[1] - 3.0, hi[1], n)   # last rows along y
    p1[:, 2] = lo[2] - 5.0
    p2 = p1.copy()
    p2[:, 0] += rng.uniform(-1.0, 1.0, n)
    p2[:, 1] += rng.uniform(-1.0, 1.0, n)
    p2[:, 2] = hi[2] + 5.0
    pc = p1 + rng.uniform(0.8, 0.97, n)[:, None] * (p2 - p1)  # TOF centres in the last slices
    f = rng.uniform(0.5, 1.5, n)
    lors = np.vstack([p1.T, p2.T, pc.T, f[None, :]]).astype(np.float32)
    img = H.image(H.TOR_GRID, seed=9)
    L, X = _t(lors, cuda), _t(img, cuda)
    args = (H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, False, H.TOF_BIN, H.TOF_SIGMA2)
    plan = _plan(L, False, True, (8, 16, 16), 256, tiled_projection=False)
    got = plan.osem_subset(X).cpu().numpy()
    ref = ops.osem_subset(L, X, *args).cpu().numpy()
    plan.close()
    assert float(np.abs(ref[-1, -1, :]).sum()) > 0  # the last row of the last slice is hit
    assert H.rel_l2(got, ref) <= 1e-6
    p = ops.projection(L, X, *args).cpu().numpy()
    pref = O.tor_project(lors, img, *args, prec="f64")
    assert H.rel_l2(p, pref) <= TOL64


# (tile, chunk): the default tile (per-workgroup counters in LDS, the fill
# pass with listing queues); small tiles whose dozens of listings per event
# overflow the fill pass's queues (2 x 4 x 4, 1080 tiles) or shorten them
# and overflow them (1 x 4 x 4, 2160 tiles: 768-entry queues); and a tile so
# small that the counters of two plans exceed the LDS (global counters)
@pytest.mark.parametrize("tile,chunk", [((8, 32, 20), 2048), ((2, 4, 4), 256),
                                        ((1, 4, 4), 512), ((1, 2, 2), 50)])
def test_plan_subsets_equal_single_plans(cuda, case, tile, chunk):
    """Plans of consecutive event ranges built together
    (srfx_tor_plan_create_subsets) equal plans built one by one: the same
    listings, the same OSEM step and backprojection (to FP32 rounding of the
    LDS and flush order), and the FP64 oracle.  The ranges include an empty
    one and ranges smaller than a workgroup (a workgroup then spans more
    than two plans)."""
    from oracle import lib as O
    from srf_amd.plan import TorPlan
    lors, img = case
    L, X = _t(lors, cuda), _t(img, cuda)
    n = lors.shape[1]
    bounds = [0, 3, 3, 100, 2500, 2600, n]
    args = (H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, True, H.TOF_BIN, H.TOF_SIGMA2)
    plans = TorPlan.subsets(L, bounds, *args, tile=tile, chunk=chunk)
    assert len(plans) == len(bounds) - 1
    v = np.random.default_rng(8).uniform(0.5, 2.0, n).astype(np.float32)
    for k, pl in enumerate(plans):
        lo, hi = bounds[k], bounds[k + 1]
        assert pl.n == hi - lo
        one = TorPlan(L[:, lo:hi].contiguous(), *args, tile=tile, chunk=chunk)
        assert pl.info() == one.info(), (k, pl.info(), one.info())
        if hi == lo:
            one.close()
            continue
        a, b = pl.osem_subset(X).cpu().numpy(), one.osem_subset(X).cpu().numpy()
        assert H.rel_l2(a, b) <= 1e-6
        sub = np.ascontiguousarray(lors[:, lo:hi])
        p = O.tor_project(sub, img, *args, prec="f64")
        keep = p > 1e-7
        inv = np.where(keep, 1.0 / np.where(keep, p, 1.0), 0.0).astype(np.float32)
        ref = O.tor_backproject_adjoint(sub, inv, *args)
        assert H.rel_l2(a, ref) <= TOL64
        bp = pl.backprojection(_t(v[lo:hi], cuda)).cpu().numpy()
        rb = O.tor_backproject(sub, v[lo:hi], *args, prec="f64")
        assert H.rel_l2(bp, rb) <= TOL64
        one.close()
    for pl in plans[::-1]:  # the shared blocks go with the last plan
        pl.close()


def test_plan_subset_groups_equal_subsets(cuda, case):
    """Subset plans of several LOR arrays built in one call
    (srfx_tor_plan_create_subset_groups), each group with its own tile and
    bounds, equal each group's plans built by srfx_tor_plan_create_subsets:
    the same listings and walks, the same OSEM step (to FP32 rounding of the
    flush order)."""
    from srf_amd.plan import TorPlan
    lors, img = case
    L, X = _t(lors, cuda), _t(img, cuda)
    n = lors.shape[1]
    L2 = torch.flip(L, dims=[1]).contiguous()  # the same events in another order
    args = (H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE)
    model = (H.KW, True, H.TOF_BIN, H.TOF_SIGMA2)
    specs = [dict(lors=L, bounds=[0, 1000, n], tile=(8, 32, 20)),
             dict(lors=L[:, :0], bounds=[0, 0, 0], tile=(8, 32, 20)),  # a group without events
             dict(lors=L2, bounds=[0, 3, 3, 2000, n], tile=(12, 32, 10), axis="y")]
    groups = [dict(s, grid=args[0], center=args[1], size=args[2]) for s in specs]
    built = TorPlan.subset_groups(groups, *model)
    assert [len(b) for b in built] == [2, 2, 4]
    assert all(pl.n == 0 and pl.info() == (0, 0) for pl in built[1])
    for spec, plans in zip(specs, built):
        ref = TorPlan.subsets(spec["lors"], spec["bounds"], *args, *model, tile=spec["tile"])
        for pl, rp in zip(plans, ref):
            assert pl.n == rp.n and pl.tile == rp.tile and pl.axis == spec.get("axis")
            assert pl.info() == rp.info() and pl.walks() == rp.walks()
            if pl.n:
                a, b = pl.osem_subset(X).cpu().numpy(), rp.osem_subset(X).cpu().numpy()
                assert H.rel_l2(a, b) <= 1e-6
        for rp in ref[::-1]:
            rp.close()
    for plans in built[::-1]:
        for pl in plans[::-1]:
            pl.close()
