"""Checks at the TOF OSEM workload's own shape: a 200 x 200 x 100 image of
2.0 x 2.0 x 2.2 mm voxels, TOF at 530 ps FWHM, 4 mm tube half-width, and
2M list-mode events (one OSEM subset of the 20M-event set), in the
locality-sorted order the reconstruction uses.  The FP64 oracle does not
finish at this size in seconds, so these checks hold at any size:

* the adjoint identity of the whole model (every axis group):
  |<Ax, y> - <x, A^T y>| / |<Ax, y>| <= 1e-5, untiled and through tile plans;
* the tiled backprojection (FP64 LDS tiles) equals the untiled one (FP32
  global atomics) to rel L2 1e-5: two summation orders of one sum;
* one OSEM subset step through tile plans (projection kernel + tiled
  scatter of 1/p) equals the fused untiled kernel to rel L2 1e-5;
* the list-mode count identity of that step: sum_j x_j bp_j equals the
  number of events with p_i > 1e-7.
"""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

N_EVENTS = 2_000_000


@pytest.fixture(scope="module")
def workload(cuda):
    from srf_amd import physics, preprocess, synth
    from srf_amd.data import Image, split_from_groups
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner()
    _, ev, meta = synth.c2(n_events=N_EVENTS, seed=21, scanner=sc, device=cuda)
    groups = preprocess.recon_process(ev, sc.tof_res_ps)
    del ev
    groups, _ = preprocess.reorder_groups(groups, 1)
    data = split_from_groups(groups)
    grid, center, size = meta["grid"], meta["center"], meta["size"]
    assert tuple(grid) == (200, 200, 100)
    model = physics.SplitLoRsModel(meta["kernel_width"], meta["geo_sigma2_flag"],
                                   tof_sigma2=meta["tof_sigma2"], tof_bin=meta["tof_bin"])
    x = torch.from_numpy(H.image(tuple(grid)[::-1], seed=9).reshape(grid)).to(cuda)
    image = Image(x, center, size)
    plans = model.make_plans(image, data)
    print(f"full-size workload: {N_EVENTS} events, groups "
          + ", ".join(f"{a}={len(data[a])}" for a in "xyz"), flush=True)
    yield model, data, image, plans
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
    model, data, image, plans = workload
    v = _ratio_values(data, cuda)
    p = model.projection(image, data)
    lhs = sum(float(torch.dot(p[a].values.double(), 1.0 / v[a].double())) for a in "xyz"
              if len(data[a]))
    bp = model.backprojection(_with_values(data, v), image,
                              plans=plans if tiled else None).data
    rhs = float(torch.sum(image.data.double() * bp.double()))
    e = abs(lhs - rhs) / abs(lhs)
    print(f"adjoint identity tiled={tiled}: <Ax,y> {lhs:.6e} <x,A^T y> {rhs:.6e} rel {e:.2e}")
    assert e <= 1e-5


def test_tiled_backprojection_equals_untiled(cuda, workload):
    model, data, image, plans = workload
    d = _with_values(data, _ratio_values(data, cuda))
    a = model.backprojection(d, image).data
    b = model.backprojection(d, image, plans=plans).data
    e = H.rel_l2(b.cpu().numpy(), a.cpu().numpy())
    print(f"tiled vs untiled backprojection rel_l2 {e:.2e}")
    assert e <= 1e-5


def test_plan_osem_step_equals_fused(cuda, workload):
    model, data, image, plans = workload
    fused = model.osem_subset(image, data).data
    tiled = model.osem_subset(image, data, plans=plans).data
    e = H.rel_l2(tiled.cpu().numpy(), fused.cpu().numpy())
    p = model.projection(image, data)
    n_pos = sum(int((p[a].values > 1e-7).sum()) for a in "xyz" if len(data[a]))
    lhs = float(torch.sum(image.data.double() * tiled.double()))
    print(f"plan OSEM step vs fused rel_l2 {e:.2e}; count identity {lhs:.6e} vs {n_pos}")
    assert e <= 1e-5
    assert lhs == pytest.approx(n_pos, rel=1e-5)
    assert np.isfinite(tiled.cpu().numpy()).all()
