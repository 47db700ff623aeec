"""Differentiable projector: the backward pass is the exact adjoint.

* torch.autograd.gradcheck-style identity: <d(p.g)/dx, e> = <A e, g> for the
  autograd gradient (checked through the adjoint identity);
* the gradient of the list-mode Poisson NLL at 64 random voxels against
  central finite differences (eps = 1e-3) evaluated in FP64 with the oracle:
  L(x +- eps e_j) uses the oracle's FP64 projections p and the oracle's FP64
  column a_.j (the projection of the unit image e_j), which is exact for the
  linear projector;
* the autograd gradient equals s - BackProjection(lors_value = p) of the
  reference-semantics kernel (A^T(1/p), the reference backprojector).
"""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

GRID = (30, 28, 20)
CENTER = (0.0, 0.0, 0.0)
SIZE = (120.0, 112.0, 88.0)   # 4.0 x 4.0 x 4.4 mm
KW = 8.0


@pytest.fixture(scope="module")
def problem(cuda):
    """One group of LORs in the group frame = world frame of the z group."""
    lors = H.tor_lors(20000, seed=31, grid=GRID, center=CENTER, size=SIZE)
    rng = np.random.default_rng(3)
    x = rng.uniform(0.5, 1.5, GRID[::-1]).astype(np.float32)  # op frame [gz][gy][gx]
    s = rng.uniform(0.5, 2.0, GRID[::-1]).astype(np.float32)
    return lors, x, s


def _project_fn(lors_t, shape):
    from srf_amd import ops

    class F(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            return ops.projection(lors_t, x, GRID, CENTER, SIZE, KW, False, H.TOF_BIN,
                                  H.TOF_SIGMA2)

        @staticmethod
        def backward(ctx, g):
            return ops.backprojection_adjoint(torch.empty(shape, device=g.device), lors_t,
                                              g.contiguous(), GRID, CENTER, SIZE, H.TOF_BIN,
                                              H.TOF_SIGMA2, KW, False)
    return F


def test_model_autograd_matches_adjoint(cuda):
    """srf_amd.autograd.project on world images: gradient of <p, g> is A^T g."""
    from srf_amd import autograd, physics, preprocess, synth
    from srf_amd.data import Image, split_from_groups
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner()
    ev, _ = synth.tof_events(20000, synth.c2_phantom(seed=3, scanner=sc), seed=4, scanner=sc)
    groups = preprocess.recon_process(ev.to(cuda), sc.tof_res_ps)
    data = split_from_groups(groups)
    tb, ts2 = sc.tof_info()
    model = physics.SplitLoRsModel(20.0, False, tof_sigma2=ts2, tof_bin=tb)
    grid, center, size = (40, 40, 24), (0.0, 0.0, 0.0), (400.0, 400.0, 132.0)
    x = torch.rand(grid, device=cuda, generator=torch.Generator(device=cuda).manual_seed(0))
    x.requires_grad_(True)
    p = autograd.project(Image(x, center, size), data, model)
    g = torch.randn(p.shape, device=cuda, generator=torch.Generator(device=cuda).manual_seed(1))
    (p * g).sum().backward()
    # <A^T g, y> == <g, A y> for a random y
    y = torch.rand(grid, device=cuda, generator=torch.Generator(device=cuda).manual_seed(2))
    with torch.no_grad():
        Ay = autograd.project(Image(y, center, size), data, model)
    lhs = float(torch.dot(x.grad.reshape(-1).double(), y.reshape(-1).double()))
    rhs = float(torch.dot(g.double(), Ay.double()))
    assert abs(lhs - rhs) <= 1e-5 * abs(rhs)


def test_poisson_gradient_vs_fp64_finite_differences(cuda, problem):
    from oracle import lib as O
    from srf_amd.autograd import poisson_nll
    lors, x, s = problem
    L = torch.from_numpy(lors).to(cuda)
    X = torch.from_numpy(x).to(cuda).requires_grad_(True)
    S = torch.from_numpy(s).to(cuda)
    F = _project_fn(L, x.shape)
    p = F.apply(X)
    keep = p > 1e-7
    loss = poisson_nll(p[keep], X, S)
    loss.backward()
    grad = X.grad.cpu().numpy()

    args = (GRID, CENTER, SIZE, KW, False, H.TOF_BIN, H.TOF_SIGMA2)
    p64 = O.tor_project(lors, x, *args, prec="f64")
    m = p64 > 1e-7
    rng = np.random.default_rng(9)
    hit = np.argwhere(O.tor_backproject_adjoint(lors, np.ones(lors.shape[1], np.float32),
                                                *args[:3], KW, False, H.TOF_BIN,
                                                H.TOF_SIGMA2) > 0)
    picks = hit[rng.choice(len(hit), 64, replace=False)]
    eps = 1e-3
    for z, y_, xx in picks:
        e = np.zeros_like(x)
        e[z, y_, xx] = 1.0
        col = O.tor_project(lors, e, *args, prec="f64")  # a_{.j}
        # L(x +- eps e_j) - L(x): linear term s_j * (+-eps), log term exact by linearity
        lp = s[z, y_, xx] * eps - np.sum(np.log(p64[m] + eps * col[m]) - np.log(p64[m]))
        lm = -s[z, y_, xx] * eps - np.sum(np.log(p64[m] - eps * col[m]) - np.log(p64[m]))
        fd = (lp - lm) / (2 * eps)
        assert grad[z, y_, xx] == pytest.approx(fd, rel=2e-4, abs=1e-4), (z, y_, xx)


def test_gradient_equals_reference_backprojector(cuda, problem):
    """grad = s - A^T(1/p), where A^T(1/v) is the reference Backprojection op."""
    from srf_amd import ops
    from srf_amd.autograd import poisson_nll
    lors, x, s = problem
    L = torch.from_numpy(lors).to(cuda)
    with torch.no_grad():
        p0 = ops.projection(L, torch.from_numpy(x).to(cuda), GRID, CENTER, SIZE, KW, False,
                            H.TOF_BIN, H.TOF_SIGMA2)
    L = L[:, p0 > 1e-7].contiguous()  # events that see the image
    X = torch.from_numpy(x).to(cuda).requires_grad_(True)
    S = torch.from_numpy(s).to(cuda)
    p = _project_fn(L, x.shape).apply(X)
    assert bool((p > 1e-7).all())
    poisson_nll(p, X, S).backward()
    with torch.no_grad():
        bp = ops.backprojection(X, L, p, GRID, CENTER, SIZE, H.TOF_BIN, H.TOF_SIGMA2, KW, False)
    ref = (S - bp).cpu().numpy()
    assert H.rel_l2(X.grad.cpu().numpy(), ref) <= 1e-5


def test_autograd_with_plans_matches_untiled(cuda):
    """The backward through tile plans gives the untiled adjoint's gradient
    (FP32 summation order aside) and passes the adjoint identity."""
    from srf_amd import autograd, physics, preprocess, synth
    from srf_amd.data import Image, split_from_groups
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner()
    ev, _ = synth.tof_events(30000, synth.c2_phantom(seed=5, scanner=sc), seed=6, scanner=sc)
    data = split_from_groups(preprocess.recon_process(ev.to(cuda), sc.tof_res_ps))
    tb, ts2 = sc.tof_info()
    model = physics.SplitLoRsModel(4.0, False, tof_sigma2=ts2, tof_bin=tb)
    grid, center, size = (100, 100, 50), (0.0, 0.0, 0.0), (400.0, 400.0, 110.0)
    x0 = torch.rand(grid, device=cuda, generator=torch.Generator(device=cuda).manual_seed(3))
    plans = model.make_plans(Image(x0, center, size), data)
    g = None
    grads = []
    for pl in (None, plans):
        x = x0.clone().requires_grad_(True)
        p = autograd.project(Image(x, center, size), data, model, plans=pl)
        if g is None:
            g = torch.randn(p.shape, device=cuda,
                            generator=torch.Generator(device=cuda).manual_seed(4))
        (p * g).sum().backward()
        grads.append(x.grad.clone())
    assert H.rel_l2(grads[1].cpu().numpy(), grads[0].cpu().numpy()) <= 1e-6
    y = torch.rand(grid, device=cuda, generator=torch.Generator(device=cuda).manual_seed(5))
    with torch.no_grad():
        Ay = autograd.project(Image(y, center, size), data, model)
    lhs = float(torch.dot(grads[1].reshape(-1).double(), y.reshape(-1).double()))
    rhs = float(torch.dot(g.double(), Ay.double()))
    assert abs(lhs - rhs) <= 1e-5 * abs(rhs)
    for pl in plans.values():
        if pl is not None:
            pl.close()
