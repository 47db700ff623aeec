"""Differentiable TOR projector.

``project(image, data, model, plans=None)`` returns the projections of every
event of ``data`` (groups x, y, z concatenated) and is differentiable with
respect to ``image``.  The backward pass is the hand-written plain adjoint
(sum_i w_ij g_i; srfx_tor_backproject_adjoint, or the tiled
srfx_tor_plan_backproject_adjoint with ``plans``): the system-matrix weights
are recomputed from the LORs, no per-event voxel lists are stored.  Only the
image's shape and the LOR tensors (already resident) are kept for backward.
TOF sigma and kernel width are attributes, not differentiated.

``poisson_nll(p, x, sensitivity)`` is the list-mode Poisson negative
log-likelihood  sum_j s_j x_j - sum_i log p_i  whose gradient
s - A^T(1/p) drives the MLEM update x <- x * A^T(1/p) / s.

``stored_list_bytes(model, image, data)`` is the memory a stored-voxel-list
backward would need (an int32 index and a float32 weight per non-zero entry),
for comparison with recomputation.
"""
import torch

from . import _lib
from .data import Image
from .ops import _ptr, _soa, _stream


def _grid(model, image, axis):
    return model._grid(image, axis)


class _TorProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, model, center, size, soas, plans):
        img = Image(x.contiguous(), center, size)
        outs = []
        for a, soa in zip(("x", "y", "z"), soas):
            lors, n, ld = _soa(soa, 10)
            p = torch.empty(n, dtype=torch.float32, device=x.device)
            if n:
                src, g = model._source(img, a)  # op-frame copy when the model stages
                _lib.call("srfx_tor_project", _ptr(lors), n, ld, _ptr(src), g, model._model(),
                          _ptr(p), model._flags("projection"), _stream())
            outs.append(p)
        ctx.model = model
        ctx.plans = plans
        ctx.center, ctx.size = center, size
        ctx.soas = soas
        ctx.shape = x.shape
        ctx.counts = [o.numel() for o in outs]
        return torch.cat(outs)

    @staticmethod
    def backward(ctx, g):
        g = g.contiguous().to(torch.float32)
        out = torch.zeros(ctx.shape, dtype=torch.float32, device=g.device)
        img = Image(out, ctx.center, ctx.size)
        off = 0
        for a, soa, n in zip(("x", "y", "z"), ctx.soas, ctx.counts):
            if n:
                gi = g[off:off + n]
                if ctx.plans is not None:
                    ctx.model._planned(ctx.plans, a, n, out).backprojection_adjoint(
                        gi, out=out, accumulate=True)
                else:
                    lors, n_, ld = _soa(soa, 10)
                    _lib.call("srfx_tor_backproject_adjoint", _ptr(lors), n_, ld, _ptr(gi),
                              _grid(ctx.model, img, a), ctx.model._model(), _ptr(out),
                              _lib.SRFX_ACCUMULATE, _stream())
            off += n
        return out, None, None, None, None, None


def project(image, data, model, plans=None):
    """Differentiable projection of Image ``image`` (data (nx, ny, nz),
    requires_grad as needed) for ListModeDataSplit ``data``.  With ``plans``
    (model.make_plans of the same events, for images laid out like ``image``)
    the backward pass scatters through the tile plans."""
    soas = tuple(data[a].soa for a in ("x", "y", "z"))
    return _TorProject.apply(image.data, model, list(image.center), list(image.size), soas,
                             plans)


def poisson_nll(p, x, sensitivity):
    """sum_j s_j x_j - sum_i log p_i (list mode, one count per event)."""
    return torch.sum(sensitivity * x) - torch.sum(torch.log(p))


def stored_list_bytes(model, image, data):
    """Bytes of an explicit (int32 voxel index, float32 weight) list for every
    non-zero system-matrix entry of ``data``."""
    return int(model.count_nnz(image, data).item()) * 8
