"""srfx_tor_osem_subset_noalloc: the fused OSEM subset pass under a slab
schedule with its per-slab partial projections in the library's static
scratch (no allocation), called through the C ABI.  It must give the results
of srfx_tor_osem_subset with the same schedule (a pool workspace) to FP32
atomic-order rounding, on two calls in a row, the second with fewer events
(the scratch keeps the first call's partials, which the second must not
read), and reject inputs beyond the scratch before any device work."""
import ctypes

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

ARGS = (H.TOR_GRID, H.TOR_CENTER, H.TOR_SIZE)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _call(name, L, X, out, flags):
    import torch
    from srf_amd import _lib
    from srf_amd.ops import _ptr, _stream
    g = _lib.make_grid(*ARGS)
    m = _lib.make_tor_model(H.KW, True, H.TOF_BIN, H.TOF_SIGMA2)
    n = L.shape[1]
    _lib.call(name, _ptr(L), n, n, _ptr(X), g, g, m, _ptr(out), flags, _stream())
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("slab", ["xcd8", "xcd16"])
def test_noalloc_equals_pool_workspace(cuda, slab):
    import torch
    from oracle import lib as O
    from srf_amd import _lib
    flags = _lib.schedule_flags((1, 256, slab))
    X = _t(H.image(H.TOR_GRID, seed=5), cuda)
    for k, n in enumerate((6000, 2500)):
        lors = H.tor_lors(n, seed=40 + k)
        L = _t(lors, cuda)
        print(f"noalloc {slab} call {k}: n={n}", flush=True)
        a = _call("srfx_tor_osem_subset_noalloc", L, X, torch.empty_like(X), flags)
        b = _call("srfx_tor_osem_subset", L, X, torch.empty_like(X), flags)
        assert float(np.abs(b).sum()) > 0
        e = H.rel_l2(a, b)
        print(f"  vs pool workspace rel_l2 {e:.2e}")
        assert e <= 1e-6
        p = O.tor_project(lors, H.image(H.TOR_GRID, seed=5), *ARGS, H.KW, True, H.TOF_BIN,
                          H.TOF_SIGMA2, prec="f64")
        keep = p > 1e-7
        inv = np.where(keep, 1.0 / np.where(keep, p, 1.0), 0.0).astype(np.float32)
        ref = O.tor_backproject_adjoint(lors, inv, *ARGS, H.KW, True, H.TOF_BIN, H.TOF_SIGMA2)
        assert H.rel_l2(a, ref) <= 1e-5


def test_noalloc_default_slabs_and_accumulate(cuda):
    """Without a slab field the entry point runs SRFX_SLAB_XCD8; with
    SRFX_ACCUMULATE it adds into bp."""
    import torch
    from srf_amd import _lib
    X = _t(H.image(H.TOR_GRID, seed=6), cuda)
    L = _t(H.tor_lors(3000, seed=44), cuda)
    ref = _call("srfx_tor_osem_subset", L, X, torch.empty_like(X),
                _lib.schedule_flags((1, 256, "xcd8")))
    got = _call("srfx_tor_osem_subset_noalloc", L, X, torch.full_like(X, 0.5),
                _lib.SRFX_ACCUMULATE)
    assert H.rel_l2(got - 0.5, ref) <= 1e-6


def test_noalloc_rejects_inputs_beyond_scratch(cuda):
    """slabs x n above the scratch returns SRFX_ETOOLARGE before any device
    work (the LOR pointer is never read)."""
    import torch
    from srf_amd import _lib
    lib = _lib.load()
    g = _lib.make_grid(*ARGS)
    m = _lib.make_tor_model(H.KW, False, H.TOF_BIN, H.TOF_SIGMA2)
    X = torch.ones(H.TOR_GRID[::-1], device=cuda)
    out = torch.zeros_like(X)
    n = _lib.NOALLOC_SCRATCH_FLOATS // 16 + 1
    fake = ctypes.c_void_p(0x1000)
    st = lib.srfx_tor_osem_subset_noalloc(fake, n, n, X.data_ptr(), g, g, m, out.data_ptr(),
                                          _lib.schedule_flags((1, 256, "xcd16")), None)
    assert st == _lib.SRFX_ETOOLARGE
    torch.cuda.synchronize()
    assert float(out.abs().sum()) == 0.0
