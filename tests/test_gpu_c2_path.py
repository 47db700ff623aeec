"""The TOF OSEM workload's own kernels at its own geometry against the FP64
oracle: a 200 x 200 x 100 image of 2.0 x 2.0 x 2.2 mm voxels, TOF 530 ps,
kernel_width 4 mm, and a seeded 150k-event sample of the 20M-event workload
(same generator and phantom), preprocessed and locality-ordered as the
benchmark does.  The path is the benchmark's exactly: per axis group the
op-frame copy of the image, the record projection (4 lanes per event,
frame gathers, record_proj_kernel) and the tiled scatter of 1/p (default
tile 12 x 32 x 10, chunk 2048, FP64 LDS tile, world-layout flush; every
event walks the 3 x 2 box, so the scatter runs its single-walk kernel).

* projection (read back from the scales the projection stores, 1/p) and
  backprojection of one OSEM step: rel L2 <= 1e-5 against the FP64 oracle;
* OSEM 2 iterations x 2 subsets: rel L2 <= 1e-4;
* the error of the reference's own FP32 arithmetic (the FP32 port) at these
  scanner-scale coordinates is printed beside it.
"""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

N_EVENTS = 150_000


@pytest.fixture(scope="module")
def workload(cuda):
    from srf_amd import physics, preprocess, synth
    from srf_amd.data import split_from_groups
    from srf_amd.plan import DEFAULT_CHUNK, DEFAULT_TILE
    from srf_amd.scanner import CylindricalScanner
    assert DEFAULT_TILE == (12, 32, 10) and DEFAULT_CHUNK == 2048
    sc = CylindricalScanner()
    _, ev, meta = synth.c2(n_events=N_EVENTS, seed=2, scanner=sc, device=cuda)
    groups = preprocess.recon_process(ev, sc.tof_res_ps)
    del ev
    groups, _ = preprocess.reorder_groups(groups, 2)
    model = physics.SplitLoRsModel(meta["kernel_width"], meta["geo_sigma2_flag"],
                                   tof_sigma2=meta["tof_sigma2"], tof_bin=meta["tof_bin"])
    model.schedule["projection"] = (4, 256)  # the benchmark's projection schedule
    assert model.stage_frames
    gnp = {a: groups[a].cpu().numpy() for a in "xyz"}
    print(f"C2 sample: {N_EVENTS} events, groups "
          + ", ".join(f"{a}={gnp[a].shape[1]}" for a in "xyz"), flush=True)
    return model, groups, gnp, meta, split_from_groups(groups)


def test_osem_step_vs_oracle(cuda, workload):
    from oracle import recon as R
    from srf_amd.data import Image
    model, groups, gnp, meta, data = workload
    grid, center, size = meta["grid"], meta["center"], meta["size"]
    kw, tb, ts2 = meta["kernel_width"], meta["tof_bin"], meta["tof_sigma2"]
    x_np = H.image(tuple(grid)[::-1], seed=21).reshape(grid)
    x = torch.from_numpy(x_np).to(cuda)
    image = Image(x, center, size)
    plans = model.make_plans(image, data)
    assert all(plans[a].walks() == (3, True) for a in "xy")
    out = torch.zeros_like(x)
    model.osem_subset(image, data, out=out, plans=plans)
    bp = out.cpu().numpy()
    p64 = R.project_split(gnp, x_np, center, size, kw, False, tb, ts2, prec="f64")
    p32 = R.project_split(gnp, x_np, center, size, kw, False, tb, ts2, prec="f32")
    pg, pr, pp = [], [], []
    for a in "xy":
        s = plans[a].scales().cpu().numpy().astype(np.float64)
        keep = s > 0
        assert np.array_equal(keep, p64[a] > 1e-7)
        pg.append(np.where(keep, 1.0 / np.where(keep, s, 1.0), 0.0))
        pr.append(np.where(keep, p64[a], 0.0))
        pp.append(np.where(keep, p32[a], 0.0))
    for p in plans.values():
        if p is not None:
            p.close()
    pg, pr, pp = np.concatenate(pg), np.concatenate(pr), np.concatenate(pp)
    e_proj = H.rel_l2(pg, pr)
    e_proj32 = H.rel_l2(pp, pr)
    v = {a: p64[a].astype(np.float32) for a in "xyz"}
    b64 = R.backproject_split(gnp, v, tuple(grid), center, size, kw, False, tb, ts2, prec="f64")
    v32 = {a: p32[a].astype(np.float32) for a in "xyz"}
    b32 = R.backproject_split(gnp, v32, tuple(grid), center, size, kw, False, tb, ts2,
                              prec="f32")
    e_bp = H.rel_l2(bp, b64)
    e_bp32 = H.rel_l2(b32, b64)
    print(f"C2 path vs FP64 oracle: projection rel_l2 {e_proj:.2e} (FP32 port {e_proj32:.2e}), "
          f"OSEM-step backprojection rel_l2 {e_bp:.2e} (FP32 port {e_bp32:.2e})")
    assert e_proj <= 1e-5
    assert e_bp <= 1e-5


def test_single_walk_scatter_matches_general(cuda, workload):
    """The single-walk scatter kernel (tile 12 x 32 x 10: five workgroups per
    CU) and the general one (tile 12 x 32 x 12, whose FP64 LDS tile admits
    four) give the same OSEM-step backprojection."""
    from srf_amd.data import Image
    model, groups, gnp, meta, data = workload
    grid, center, size = meta["grid"], meta["center"], meta["size"]
    x = torch.from_numpy(H.image(tuple(grid)[::-1], seed=23).reshape(grid)).to(cuda)
    image = Image(x, center, size)
    outs = []
    try:
        for tile in ((12, 32, 10), (12, 32, 12)):
            model.tile = tile
            plans = model.make_plans(image, data)
            assert all(plans[a].walks() == (3, True) for a in "xy")
            out = torch.zeros_like(x)
            model.osem_subset(image, data, out=out, plans=plans)
            outs.append(out.cpu().numpy())
            for p in plans.values():
                if p is not None:
                    p.close()
    finally:
        model.tile = None
    e = H.rel_l2(outs[0], outs[1])
    print(f"single-walk vs general scatter kernel: rel_l2 {e:.2e}")
    assert e <= 1e-6


def test_osem_2x2_vs_oracle(cuda, workload):
    from oracle import recon as R
    from srf_amd.data import Image
    from srf_amd.recon import OSEM
    model, groups, gnp, meta, data = workload
    grid, center, size = meta["grid"], meta["center"], meta["size"]
    kw, tb, ts2 = meta["kernel_width"], meta["tof_bin"], meta["tof_sigma2"]
    rng = np.random.default_rng(22)
    emap = rng.uniform(0.5, 1.5, grid).astype(np.float32)
    x0 = np.ones(grid, np.float32)
    x = torch.from_numpy(x0).to(cuda)
    osem = OSEM(model, data, Image(torch.from_numpy(emap).to(cuda), center, size),
                nb_subsets=2, tiled=True)
    osem.iterate(Image(x, center, size), 2)
    got = x.cpu().numpy()
    osem.close()
    ref = R.osem(gnp, x0, emap.astype(np.float64), center, size, kw, False, tb, ts2, 2, 2)
    e = H.rel_l2(got, ref)
    print(f"C2 path OSEM 2 it x 2 subsets vs FP64 oracle rel_l2 {e:.2e}")
    assert e <= 1e-4
