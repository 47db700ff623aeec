"""Checks at full image sizes, where the FP64 oracle does not finish in
seconds; they hold at any size.  Workloads:

* c2: 200 x 200 x 100 of 2.0 x 2.0 x 2.2 mm, TOF 530 ps, 2M events (one
  OSEM subset of the 20M-event set), locality-sorted as the reconstruction
  uses them;
* c5: 256 x 256 x 160 of 1.6 x 1.6 x 1.7 mm (candidate box 4), TOF 400 ps,
  2M events from the torso phantom;
* c4_200, c4_128: members of the batched workload, 200 x 200 x 100 and
  128 x 128 x 64 of 3.0 x 3.0 x 3.4 mm, 1M events each.

All at kernel_width 4 mm.  Per workload:

* the adjoint identity of the whole model (every axis group):
  |<Ax, y> - <x, A^T y>| / |<Ax, y>| <= 1e-5, untiled and through tile plans;
* the tiled backprojection (FP64 LDS tiles) equals the untiled one (FP32
  global atomics) to rel L2 1e-5: two summation orders of one sum;
* one OSEM subset step through tile plans (projection kernel + tiled
  scatter of 1/p) equals the fused untiled kernel to rel L2 1e-5;
* the list-mode count identity of that step: sum_j x_j bp_j equals the
  number of events with p_i > 1e-7.

And for c5 the attenuated efficiency map (Maplors with weights
1 / exp(-int mu), mu 0.0096/mm in the body) over 3M crystal pairs: tiled
equals untiled (rel L2 <= 1e-5), and voxel by voxel it lies between
exp(-mu 360 mm) and 1 times the unattenuated map.
"""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

WORKLOADS = ["c2", "c5", "c4_200", "c4_128"]


def _make(name, cuda):
    from srf_amd import preprocess, synth
    from srf_amd.scanner import CylindricalScanner
    if name == "c2":
        sc = CylindricalScanner()
        _, ev, meta = synth.c2(n_events=2_000_000, seed=21, scanner=sc, device=cuda)
        grid, center, size = meta["grid"], meta["center"], meta["size"]
        extra = {}
    elif name == "c5":
        sc = CylindricalScanner(tof_res_ps=400.0)
        grid, center, size = (256, 256, 160), (0.0, 0.0, 0.0), (409.6, 409.6, 272.0)
        act, mu = synth.torso_image(grid, center, size, seed=5, device=cuda)
        ev, _ = synth.tof_events_from_image(2_000_000, act, center, size, seed=500, scanner=sc,
                                            device=cuda)
        extra = {"mu": mu}
    else:
        sc = CylindricalScanner()
        if name == "c4_200":
            grid, center, size, n = (200, 200, 100), (0.0, 0.0, 0.0), (400.0, 400.0, 220.0), 10**6
        else:
            grid, center, size, n = (128, 128, 64), (0.0, 0.0, 0.0), (384.0, 384.0, 217.6), 10**6
        ev, _ = synth.tof_events(n, synth.c2_phantom(seed=300, scanner=sc), seed=400,
                                 scanner=sc, device=cuda)
        extra = {}
    groups = preprocess.recon_process(ev, sc.tof_res_ps)
    del ev
    groups, _ = preprocess.reorder_groups(groups, 1)
    return sc, tuple(grid), tuple(center), tuple(size), groups, extra


@pytest.fixture(scope="module", params=WORKLOADS)
def workload(request, cuda):
    from srf_amd import _lib, physics
    from srf_amd.data import Image, split_from_groups
    sc, grid, center, size, groups, extra = _make(request.param, cuda)
    data = split_from_groups(groups)
    tb, ts2 = sc.tof_info()
    model = physics.SplitLoRsModel(4.0, False, tof_sigma2=ts2, tof_bin=tb)
    model.schedule["projection"] = (4, 256)
    x = torch.from_numpy(H.image(grid[::-1], seed=9).reshape(grid)).to(cuda)
    image = Image(x, center, size)
    plans = model.make_plans(image, data)
    g, _, s, _ = physics.group_frame(grid, center, size, (1, 1, 1), "x")
    box = _lib.tor_patch_info(g, s, 4.0)[1]
    print(f"full-size workload {request.param} {grid}: box {box}, groups "
          + ", ".join(f"{a}={len(data[a])}" for a in "xyz"), flush=True)
    yield request.param, model, data, image, plans, sc, extra
    for p in plans.values():
        if p is not None:
            p.close()


def _with_values(data, vals):
    from srf_amd.data import ListModeDataSplit
    return ListModeDataSplit(*(data[a].replace_values(vals[a]) for a in "xyz"))


def _ratio_values(data, cuda):
    """Per group v in [0.5, 2): the backprojection adds A^T (1/v)."""
    g = torch.Generator(device=cuda).manual_seed(3)
    return {a: 0.5 + 1.5 * torch.rand(len(data[a]), generator=g, device=cuda) for a in "xyz"}


@pytest.mark.parametrize("tiled", [False, True])
def test_adjoint_identity(cuda, workload, tiled):
    name, model, data, image, plans, _, _ = workload
    v = _ratio_values(data, cuda)
    p = model.projection(image, data)
    lhs = sum(float(torch.dot(p[a].values.double(), 1.0 / v[a].double())) for a in "xyz"
              if len(data[a]))
    bp = model.backprojection(_with_values(data, v), image,
                              plans=plans if tiled else None).data
    rhs = float(torch.sum(image.data.double() * bp.double()))
    e = abs(lhs - rhs) / abs(lhs)
    print(f"{name} adjoint identity tiled={tiled}: <Ax,y> {lhs:.6e} <x,A^T y> {rhs:.6e} "
          f"rel {e:.2e}")
    assert e <= 1e-5


def test_tiled_backprojection_equals_untiled(cuda, workload):
    name, model, data, image, plans, _, _ = workload
    d = _with_values(data, _ratio_values(data, cuda))
    a = model.backprojection(d, image).data
    b = model.backprojection(d, image, plans=plans).data
    e = H.rel_l2(b.cpu().numpy(), a.cpu().numpy())
    print(f"{name} tiled vs untiled backprojection rel_l2 {e:.2e}")
    assert e <= 1e-5


def test_plan_osem_step_equals_fused(cuda, workload):
    name, model, data, image, plans, _, _ = workload
    fused = model.osem_subset(image, data).data
    tiled = model.osem_subset(image, data, plans=plans).data
    e = H.rel_l2(tiled.cpu().numpy(), fused.cpu().numpy())
    p = model.projection(image, data)
    n_pos = sum(int((p[a].values > 1e-7).sum()) for a in "xyz" if len(data[a]))
    lhs = float(torch.sum(image.data.double() * tiled.double()))
    print(f"{name} plan OSEM step vs fused rel_l2 {e:.2e}; count identity {lhs:.6e} vs {n_pos}")
    assert e <= 1e-5
    assert lhs == pytest.approx(n_pos, rel=1e-5)
    assert np.isfinite(tiled.cpu().numpy()).all()


def test_attenuated_map_c5(cuda):
    from srf_amd import attenuation, effmap, physics, synth
    from srf_amd.data import Image
    from srf_amd.scanner import CylindricalScanner
    name = "c5"
    sc = CylindricalScanner(tof_res_ps=400.0)
    grid, center, size = (256, 256, 160), (0.0, 0.0, 0.0), (409.6, 409.6, 272.0)
    _, mu = synth.torso_image(grid, center, size, seed=5, device=cuda)
    tb, ts2 = sc.tof_info()
    model = physics.SplitLoRsModel(4.0, False, tof_sigma2=ts2, tof_bin=tb)
    mu = Image(mu, center, size)
    kw = dict(max_pairs=3_000_000, chunk=1_500_000)
    a = attenuation.attenuated_tor_effmap(sc, model, mu, grid, center, size, cuda, **kw)
    b = attenuation.attenuated_tor_effmap(sc, model, mu, grid, center, size, cuda, tiled=True,
                                          **kw)
    m0 = effmap.compute_tor_effmap(sc, model, grid, center, size, cuda, max_pairs=3_000_000,
                                   chunk=1_500_000)
    e = H.rel_l2(b.cpu().numpy(), a.cpu().numpy())
    an, m0n = a.double().cpu().numpy(), m0.double().cpu().numpy()
    lo = np.exp(-0.0096 * 360.0)
    print(f"{name} attenuated map, 3M pairs: tiled vs untiled rel_l2 {e:.2e}, "
          f"sum {an.sum():.6e} of {m0n.sum():.6e} unattenuated")
    assert e <= 1e-5
    assert m0n.sum() > 0
    assert np.all(an <= m0n * (1 + 1e-5) + 1e-12)
    assert np.all(an >= m0n * lo * (1 - 1e-5) - 1e-12)
