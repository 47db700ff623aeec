"""Edge cases of the TOR and Siddon ops on the GPU against the FP64 oracle:
empty inputs, ragged event counts under every launch schedule, and events
that take the rare exact paths or none at all (crossings exactly on voxel
boundaries, end points coinciding, no crossed slice, NaN rows, events far
outside the image), through the untiled kernels and the tile plans."""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def test_empty_inputs(cuda):
    import torch
    from srf_amd import ops
    g, c, s = H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE
    img = torch.ones(g[::-1], device=cuda)
    l10, l7 = torch.zeros(10, 0, device=cuda), torch.zeros(7, 0, device=cuda)
    v = torch.zeros(0, device=cuda)
    assert ops.projection(l10, img, g, c, s, H.KW).numel() == 0
    assert float(ops.backprojection(img, l10, v, g, c, s, kernel_width=H.KW).abs().sum()) == 0
    assert float(ops.backprojection_adjoint(img, l10, v, g, c, s, kernel_width=H.KW)
                 .abs().sum()) == 0
    assert float(ops.maplors(img, g, c, s, l7, v, H.KW).abs().sum()) == 0
    out = torch.full_like(img, 3.0)
    ops.osem_subset(l10, img, g, c, s, H.KW, out=out)
    assert float(out.abs().sum()) == 0  # zeroed like the reference's memset
    out.fill_(3.0)
    ops.osem_subset(l10, img, g, c, s, H.KW, out=out, accumulate=True)
    assert float(out.min()) == 3.0 and float(out.max()) == 3.0
    per, tot = ops.count_nnz(l10, g, c, s, H.KW)
    assert per.numel() == 0 and int(tot.item()) == 0
    sg, sc, ss = H.SID_GRID, H.SID_CENTER, H.SID_SIZE
    simg = torch.ones(sg[::-1], device=cuda)
    assert ops.siddon_projection(l7, simg, sg, sc, ss).numel() == 0
    assert float(ops.siddon_backprojection(simg, l7, v, sg, sc, ss).abs().sum()) == 0
    assert float(ops.siddon_maplors(simg, sg, sc, ss, l7, v).abs().sum()) == 0


@pytest.mark.parametrize("n", [1, 3, 63, 65, 257])
@pytest.mark.parametrize("sch", [None, (4, 128), (16, 64), (1, 64)])
def test_ragged_counts(cuda, n, sch):
    from oracle import lib as O
    from srf_amd import ops
    lors = H.tor_lors(n, seed=n)
    img = H.image(H.TOR_GRID, seed=2)
    args = (H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, False, H.TOF_BIN, H.TOF_SIGMA2)
    p = ops.projection(_t(lors, cuda), _t(img, cuda), *args, schedule=sch).cpu().numpy()
    ref = O.tor_project(lors, img, *args, prec="f64")
    assert p.shape == (n,)
    assert H.rel_l2(p, ref) <= 1e-5 if np.any(ref) else not np.any(p)
    v = np.random.default_rng(n).uniform(0.5, 2.0, n).astype(np.float32)
    b = ops.backprojection(_t(img, cuda), _t(lors, cuda), _t(v, cuda), H.TOR_GRID, H.TOR_CENTER,
                           H.TOR_SIZE, H.TOF_BIN, H.TOF_SIGMA2, kernel_width=H.KW,
                           schedule=sch).cpu().numpy()
    rb = O.tor_backproject(lors, v, *args, prec="f64")
    assert H.rel_l2(b, rb) <= 1e-5 if np.any(rb) else not np.any(b)


def _special_events():
    """[10, N] op-frame events: along the slice axis exactly through voxel
    boundaries (every crossing on the exact path), oblique ones that meet
    boundaries at some slices, coinciding end points, no crossed slice,
    NaN, far outside; plus ordinary events."""
    gx, gy, gz = H.TOR_GRID
    c, s = np.array(H.TOR_CENTER), np.array(H.TOR_SIZE)
    lo, hi = c - s / 2, c + s / 2
    vx = s[0] / gx
    rows = []
    for k in range(3, gx - 3, 5):
        for j in range(3, gy - 3, 7):
            x, y = lo[0] + k * vx, lo[1] + j * vx  # exact voxel boundaries
            rows.append([x, y, lo[2] - 20, x, y, hi[2] + 20])
            # oblique: dx = one voxel over one slice pitch step of 10 slices
            rows.append([x, y, lo[2] - 20, x + 4 * vx, y - 2 * vx, hi[2] + 20])
    rows.append([1.0, 2.0, 3.0, 1.0, 2.0, 3.0])                      # p1 == p2
    rows.append([lo[0], 0.0, 5.0, hi[0], 0.0, 5.0])                  # d2 = 0
    rows.append([np.nan, 0.0, lo[2] - 5, 0.0, 0.0, hi[2] + 5])       # NaN
    rows.append([1e4, 1e4, lo[2] - 5, 1e4 + 3, 1e4, hi[2] + 5])      # far outside
    rows.append([0.0, 0.0, hi[2] + 5, 0.0, 0.0, hi[2] + 40])         # beyond the last slice
    p = np.array(rows, np.float64)
    pc = 0.5 * (p[:, 0:3] + p[:, 3:6])
    f = np.ones((len(p), 1))
    special = np.hstack([p, pc, f]).T.astype(np.float32)
    return np.hstack([special, H.tor_lors(200, seed=11)])


@pytest.mark.parametrize("tof", [True, False])
def test_exact_path_and_degenerate_events(cuda, tof):
    import torch
    from oracle import lib as O
    from srf_amd import ops
    from srf_amd.plan import TorPlan
    lors = _special_events()
    n = lors.shape[1]
    img = H.image(H.TOR_GRID, seed=4)
    ts2 = H.TOF_SIGMA2 if tof else 1.0e6
    args = (H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, True, H.TOF_BIN, ts2)
    L, X = _t(lors, cuda), _t(img, cuda)
    p = ops.projection(L, X, *args).cpu().numpy()
    ref = O.tor_project(lors, img, *args, prec="f64")
    assert np.all(np.isfinite(p))
    e = H.rel_l2(p, ref)
    print(f"special events tof={tof}: projection rel_l2 vs f64 {e:.2e}")
    assert e <= 1e-5
    nd = 5  # the degenerate rows before the ordinary events
    deg = slice(n - 200 - nd, n - 200)
    assert np.all(p[deg] == 0) and np.all(ref[deg] == 0)
    per, _ = ops.count_nnz(L, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, True)
    np.testing.assert_array_equal(per.cpu().numpy(),
                                  O.tor_count(lors, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW,
                                              True))
    # the tiled scatter and the record projection take the exact paths from
    # the walk records
    # OSEM step: backprojection of 1 / p over p > 1e-7 (the oracle applies
    # the threshold to the values it is given)
    rb = O.tor_backproject(lors, ref.astype(np.float32), *args, prec="f64")
    for tile, chunk in (((8, 16, 32), 2048), ((3, 5, 7), 97)):
        plan = TorPlan(L, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, True, H.TOF_BIN, ts2,
                       tile=tile, chunk=chunk)
        b = plan.osem_subset(X).cpu().numpy()
        plan.close()
        assert np.all(np.isfinite(b))
        eb = H.rel_l2(b, rb)
        print(f"special events tof={tof} tile={tile}: osem backprojection rel_l2 {eb:.2e}")
        assert eb <= 1e-5
    b = ops.osem_subset(L, X, H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE, H.KW, True, H.TOF_BIN, ts2)
    assert torch.isfinite(b).all()
    assert H.rel_l2(b.cpu().numpy(), rb) <= 1e-5
