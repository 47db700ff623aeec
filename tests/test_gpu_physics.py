"""End-to-end GPU checks of the physics and reconstruction layers:

* strided group frames (no transposes) equal the reference's transpose-based
  dispatch (split_lors.py:77-113) bit for bit for projections;
* raw events -> product preprocessing -> SplitLoRsModel projection matches
  the oracle (oracle preprocessing + FP64 projection per group) to 1e-5;
* OSEM after 2 iterations x 3 subsets matches the oracle's MLEM/OSEM to
  rel L2 <= 1e-4 (the north star's OSEM bar), fused and unfused;
* the list-mode count identity of one update: sum_j x_j bp_j equals the number
  of events with p_i > 1e-7;
* a 90-degree rotation about z maps x-group events onto y-group events with
  identical projections of the rotated image;
* CompleteLoRsModel (Siddon) equals the oracle on the reference's transposed
  frame.
"""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

GRID = (40, 40, 24)
CENTER = (0.0, 0.0, 0.0)
SIZE = (400.0, 400.0, 132.0)  # 10 x 10 x 5.5 mm voxels
KW = 20.0                     # keeps the tube a few voxels wide at this pitch


@pytest.fixture(scope="module")
def events():
    from srf_amd import synth
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner()
    ph = synth.c2_phantom(seed=5, scanner=sc)
    ev, _ = synth.tof_events(30000, ph, seed=7, scanner=sc, device="cpu")
    return ev.numpy(), sc


def _model(sc):
    from srf_amd import physics
    tb, ts2 = sc.tof_info()
    return physics.SplitLoRsModel(KW, True, tof_sigma2=ts2, tof_bin=tb), tb, ts2


def test_strided_frames_equal_transposed_dispatch(cuda, events):
    from srf_amd import ops, preprocess
    from srf_amd.data import Image, split_from_groups
    ev, sc = events
    model, tb, ts2 = _model(sc)
    groups = preprocess.recon_process(torch.from_numpy(ev).to(cuda), sc.tof_res_ps)
    data = split_from_groups(groups)
    x = torch.rand(GRID, generator=torch.Generator().manual_seed(0)).to(cuda)
    got = model.projection(Image(x, CENTER, SIZE), data)
    perm = {"z": [2, 1, 0], "y": [1, 2, 0], "x": [0, 2, 1]}
    for a in "xyz":
        if len(data[a]) == 0:
            continue
        xa = x.permute(*perm[a]).contiguous()
        g = list(xa.shape[::-1])
        c = [CENTER[k] for k in perm[a]][::-1]
        s = [SIZE[k] for k in perm[a]][::-1]
        ref = ops.projection(data[a].soa, xa, g, c, s, KW, True, tb, ts2)
        assert torch.equal(got[a].values, ref), a


def test_pipeline_projection_vs_oracle(cuda, events):
    from oracle import preprocess as OP, recon as OR
    from srf_amd import preprocess
    from srf_amd.data import Image, split_from_groups
    ev, sc = events
    model, tb, ts2 = _model(sc)
    groups = preprocess.recon_process(torch.from_numpy(ev).to(cuda), sc.tof_res_ps)
    x = H.image(GRID[::-1], seed=3).reshape(GRID)  # any positive (nx, ny, nz) image
    p = model.projection(Image(torch.from_numpy(x).to(cuda), CENTER, SIZE),
                         split_from_groups(groups))
    ref_groups = {a: np.ascontiguousarray(g.T.astype(np.float32))
                  for a, g in OP.recon_process(ev, sc.tof_res_ps).items()}
    ref = OR.project_split(ref_groups, x, CENTER, SIZE, KW, True, tb, ts2)
    for a in "xy":
        np.testing.assert_array_equal(groups[a].cpu().numpy(), ref_groups[a])
        e = H.rel_l2(p[a].values.cpu().numpy(), ref[a])
        print(f"pipeline projection group {a}: rel_l2 vs f64 {e:.2e}")
        assert e <= 1e-5


@pytest.mark.parametrize("fused,tiled", [(True, False), (False, False), (True, True),
                                         (False, True)])
def test_osem_vs_oracle(cuda, events, fused, tiled):
    from oracle import preprocess as OP, recon as OR
    from srf_amd import preprocess
    from srf_amd.data import Image, split_from_groups
    from srf_amd.recon import OSEM
    ev, sc = events
    model, tb, ts2 = _model(sc)
    groups = preprocess.recon_process(torch.from_numpy(ev).to(cuda), sc.tof_res_ps)
    rng = np.random.default_rng(1)
    E = rng.uniform(0.5, 1.5, GRID).astype(np.float32)
    x0 = np.ones(GRID, np.float32)
    x = torch.from_numpy(x0).to(cuda)
    osem = OSEM(model, split_from_groups(groups), Image(torch.from_numpy(E).to(cuda), CENTER, SIZE),
                nb_subsets=3, fused=fused, tiled=tiled)
    if tiled:
        model.tile, model.chunk = (4, 8, 8), 300  # many tiles and chunks at this size
    osem.iterate(Image(x, CENTER, SIZE), 2)
    if tiled:
        # the same with the projection reading the world image in place
        # instead of op-frame copies
        model.stage_frames = False
        x2 = torch.from_numpy(x0).to(cuda)
        osem2 = OSEM(model, split_from_groups(groups),
                     Image(torch.from_numpy(E).to(cuda), CENTER, SIZE), nb_subsets=3,
                     fused=fused, tiled=True)
        osem2.iterate(Image(x2, CENTER, SIZE), 2)
        osem2.close()
        assert H.rel_l2(x2.cpu().numpy(), x.cpu().numpy()) <= 1e-6
    if tiled:
        assert osem.plans is not None and all(len(ps) == 3 for ps in osem.plans)
        osem.close()
    ref_groups = {a: np.ascontiguousarray(g.T.astype(np.float32))
                  for a, g in OP.recon_process(ev, sc.tof_res_ps).items()}
    ref = OR.osem(ref_groups, x0, E.astype(np.float64), CENTER, SIZE, KW, True, tb, ts2, 2, 3)
    e = H.rel_l2(x.cpu().numpy(), ref)
    print(f"OSEM 2 it x 3 subsets fused={fused} tiled={tiled}: rel_l2 vs f64 {e:.2e}")
    assert e <= 1e-4


def test_count_identity(cuda, events):
    from srf_amd import preprocess
    from srf_amd.data import Image, split_from_groups
    ev, sc = events
    model, _, _ = _model(sc)
    groups = preprocess.recon_process(torch.from_numpy(ev).to(cuda), sc.tof_res_ps)
    data = split_from_groups(groups)
    x = torch.from_numpy(H.image(GRID[::-1], seed=8).reshape(GRID)).to(cuda)
    img = Image(x, CENTER, SIZE)
    bp = model.osem_subset(img, data).data
    p = model.projection(img, data)
    n_pos = sum(int((p[a].values > 1e-7).sum()) for a in "xyz")
    lhs = float(torch.sum(x.double() * bp.double()))
    assert lhs == pytest.approx(n_pos, rel=1e-5)


def test_rotation_symmetry(cuda, events):
    """Rotate events by +90 deg about z: (x, y) -> (-y, x).  x-group events
    become y-group events; with the image rotated the same way, every event's
    projection is unchanged."""
    from srf_amd import preprocess
    from srf_amd.data import Image, split_from_groups
    ev, sc = events
    model, _, _ = _model(sc)
    ev = ev[:5000]
    rot = ev.copy()
    rot[:, 0], rot[:, 1] = -ev[:, 1], ev[:, 0]
    rot[:, 3], rot[:, 4] = -ev[:, 4], ev[:, 3]
    g0, o0 = preprocess.recon_process(torch.from_numpy(ev).to(cuda), sc.tof_res_ps,
                                      return_order=True)
    g1, o1 = preprocess.recon_process(torch.from_numpy(rot).to(cuda), sc.tof_res_ps,
                                      return_order=True)
    x = torch.from_numpy(H.image(GRID[::-1], seed=4).reshape(GRID)).to(cuda)
    # image rotated: x_rot[i, j] = x[j, n-1-i] maps world (x, y) -> (-y, x)
    xr = torch.rot90(x, 1, dims=(0, 1)).contiguous()
    p0 = model.projection(Image(x, CENTER, SIZE), split_from_groups(g0))
    p1 = model.projection(Image(xr, CENTER, SIZE), split_from_groups(g1))
    v0 = torch.zeros(len(ev), device=cuda)
    v1 = torch.zeros(len(ev), device=cuda)
    for a in "xyz":
        v0[o0[a]] = p0[a].values
        v1[o1[a]] = p1[a].values
    # ties |dx| == |dy| change groups under rotation; compare the rest
    d = np.abs(ev[:, 3:6] - ev[:, 0:3])
    keep = torch.from_numpy(d[:, 0] != d[:, 1]).to(cuda)
    e = H.rel_l2(v1[keep].cpu().numpy(), v0[keep].cpu().numpy())
    print(f"rotation symmetry rel_l2 {e:.2e}")
    assert e <= 1e-5


def test_siddon_model_vs_oracle(cuda):
    from oracle import lib as O
    from srf_amd import physics
    from srf_amd.data import Image, ListModeData
    grid, center, size = (24, 20, 16), (0.5, -1.0, 0.25), (72.0, 60.0, 54.4)
    lors = H.siddon_lors(3000, seed=9, grid=grid, center=center, size=size)
    x = H.image(grid[::-1], seed=2).reshape(grid)  # world (nx, ny, nz)
    m = physics.CompleteLoRsModel(tof_sigma2=400.0, tof_bin=6.0)
    p = m.projection(Image(torch.from_numpy(x).to(cuda), center, size),
                     ListModeData(torch.from_numpy(lors.T.copy()).to(cuda)))
    # reference frame: transpose [2,1,0] and pass grid/center/size reversed twice
    xt = np.ascontiguousarray(np.transpose(x, [2, 1, 0]))
    ref = O.siddon_project(lors, xt, grid, center, size, 6.0, 400.0, prec="f64")
    assert H.rel_l2(p.values.cpu().numpy(), ref) <= 1e-5


@pytest.mark.parametrize("shape", [(40, 36, 24), (33, 65, 7), (1, 5, 64), (200, 200, 100)])
def test_frame_copy_equals_transpose(cuda, shape):
    """srfx_frame_copy equals the reference's transpose into each group frame
    (split_lors.py:79, torch permute + contiguous), also from a strided view."""
    from srf_amd import physics
    x = torch.rand(shape, generator=torch.Generator().manual_seed(1)).to(cuda)
    for a in "xyz":
        want = x.permute(*physics.PERM[a]).contiguous()
        assert torch.equal(physics.SplitLoRsModel.frame_copy(x, a), want), a
    big = torch.rand((shape[0] + 3, shape[1] + 2, shape[2] + 5), device=cuda)
    view = big[1:1 + shape[0], 2:2 + shape[1], 0:shape[2]]
    for a in "xy":
        assert torch.equal(physics.SplitLoRsModel.frame_copy(view, a),
                           view.permute(*physics.PERM[a]).contiguous()), a
