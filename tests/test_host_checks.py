"""Host-side checks that need no GPU: the torch wrappers reject a bad
caller-supplied output image before anything is launched, and the C ABI's
patch query reproduces the reference's patch width (tor.cu.cc:375) and picks
the candidate box the kernels compile."""
import math

import pytest
import torch

from tests import helpers as H


@pytest.fixture
def host_ops(monkeypatch):
    """srf_amd.ops with the device check relaxed to dtype only (so CPU
    tensors reach the output check) and every library call recorded."""
    from srf_amd import _lib, ops
    calls = []

    def check(t, name):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch tensor")
        if t.dtype != torch.float32:
            raise TypeError(f"{name} must be float32 (got {t.dtype})")

    monkeypatch.setattr(ops, "_check_dev", check)
    monkeypatch.setattr(_lib, "call", lambda *a: calls.append(a[0]))
    monkeypatch.setattr(ops, "_stream", lambda: None)
    return ops, calls


def _bad_outs(shape):
    n = shape[0] * shape[1] * shape[2]
    return [
        (torch.zeros(n - 1), ValueError),                            # too few voxels
        (torch.zeros(n + 5), ValueError),                            # too many
        (torch.zeros(shape[::-1], dtype=torch.float16), TypeError),  # wrong dtype
        (torch.zeros(shape[::-1], dtype=torch.bfloat16), TypeError),
        (torch.zeros(shape[::-1]).transpose(0, 2), ValueError),      # not contiguous
        (torch.zeros(shape[::-1], dtype=torch.float64), TypeError),
    ]


def test_tor_osem_subset_rejects_bad_out(host_ops):
    ops, calls = host_ops
    g = H.TOR_GRID
    lors = torch.from_numpy(H.tor_lors(16, seed=1))
    img = torch.ones(g[::-1])
    for out, exc in _bad_outs(g):
        with pytest.raises(exc):
            ops.osem_subset(lors, img, g, H.TOR_CENTER, H.TOR_SIZE, H.KW, out=out)
    assert calls == []
    ops.osem_subset(lors, img, g, H.TOR_CENTER, H.TOR_SIZE, H.KW, out=torch.zeros(g[::-1]))
    assert calls == ["srfx_tor_osem_subset"]


def test_siddon_osem_subset_rejects_bad_out(host_ops):
    ops, calls = host_ops
    g = H.SID_GRID
    lors = torch.from_numpy(H.siddon_lors(16, seed=1, special=False))
    img = torch.ones(g[::-1])
    for out, exc in _bad_outs(g):
        with pytest.raises(exc):
            ops.siddon_osem_subset(lors, img, g, H.SID_CENTER, H.SID_SIZE, out=out)
    assert calls == []
    ops.siddon_osem_subset(lors, img, g, H.SID_CENTER, H.SID_SIZE, out=torch.zeros(g[::-1]))
    assert calls == ["srfx_siddon_osem_subset"]


@pytest.mark.parametrize("pitch,kw", [((2.0, 2.0, 2.2), 4.0), ((1.6, 1.6, 1.7), 4.0),
                                      ((3.0, 3.0, 3.4), 4.0), ((4.0, 4.0, 4.5), 4.0),
                                      ((2.0, 2.0, 2.2), 8.0), ((1.0, 1.0, 1.0), 8.0)])
def test_patch_info(pitch, kw):
    """P = ceil((sqrt(2) kw + inter_z) / inter_x) as the reference's host code
    forms it; the box holds the ellipse of an event along its slice axis:
    half extents sqrt(2 * 9 sigma^2) with sigma = kw / 6."""
    from srf_amd import _lib
    grid = (40, 36, 24)
    size = tuple(g * p for g, p in zip(grid, pitch))
    P, box = _lib.tor_patch_info(grid, size, kw)
    inter = [float(torch.tensor(s, dtype=torch.float32)) / g for s, g in zip(size, grid)]
    assert P == math.ceil((math.sqrt(2) * kw + inter[2]) / inter[0])
    h = math.sqrt(2 * 9 * (kw * kw / 36))
    n = max(math.floor(2 * h / inter[0]) + 1, math.floor(2 * h / inter[1]) + 1)
    assert box == (3 if n <= 3 else (4 if n <= 4 else 0))
    if pitch == (1.6, 1.6, 1.7):
        assert (P, box) == (5, 4)
    if pitch == (2.0, 2.0, 2.2) and kw == 4.0:
        assert (P, box) == (4, 3)


def test_patch_info_unsupported():
    from srf_amd import _lib
    with pytest.raises(_lib.SrfxError) as ei:
        _lib.tor_patch_info((40, 36, 24), (4.0, 3.6, 2.4), 40.0)  # P far above 16
    assert ei.value.status == _lib.SRFX_EUNSUPPORTED


def test_default_tile_by_walk():
    """Candidate-box walks take DEFAULT_TILE (the single-walk scatter fits
    five workgroups per CU), chord walks the deeper CHORD_TILE."""
    from srf_amd import _lib
    from srf_amd.plan import CHORD_TILE, DEFAULT_TILE, default_tile
    grid, size = (200, 100, 200), (400.0, 220.0, 400.0)  # the C2 y-group frame
    assert _lib.tor_patch_info(grid, size, 4.0)[1] == 3
    assert default_tile(grid, size, 4.0) == DEFAULT_TILE == (12, 32, 10)
    assert _lib.tor_patch_info(grid, size, 8.0)[1] == 0
    assert default_tile(grid, size, 8.0) == CHORD_TILE == (12, 32, 12)
