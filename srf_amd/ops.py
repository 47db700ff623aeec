"""Torch-facing wrappers of the HIP projector ops, keeping the reference op
signatures (argument names, order and meaning) of the TensorFlow custom ops:

  TOR    (src/cpp/tf_custom_ops/tor.cc:14-64)
    projection(lors, image, grid, center, size, kernel_width, geo_sigma2_flag,
               tof_bin, tof_sigma2)                                 -> [N]
    backprojection(image, lors, lors_value, grid, center, size, tof_bin,
                   tof_sigma2, kernel_width, geo_sigma2_flag)        -> image
    maplors(image, grid, center, size, lors, lors_value, kernel_width,
            geo_sigma2_flag)                                         -> image
  Siddon (src/cpp/tf_custom_ops/siddon.cc:14-57)
    siddon_projection / siddon_backprojection / siddon_maplors, same pattern.

``lors`` is the [ncols, N] structure-of-arrays tensor the reference ops take
(the physics layer passes ``transpose(lors)``, split_lors.py:81).  ``image`` is a
float32 device tensor laid out [gz][gy][gx] in the op's frame; ``grid`` is
(gx, gy, gz) and ``size`` the full box extent in mm.  ``grid``, ``center`` and
``size`` are host values (lists, tuples, numpy); the reference copies them back
from the device on every call, here they travel by value.

Every op runs on the current torch stream, on the GPU, through libsrfx.so.
"""
import torch

from . import _lib

__all__ = [
    "projection", "backprojection", "backprojection_adjoint", "maplors", "osem_subset",
    "count_nnz", "siddon_projection", "siddon_backprojection", "siddon_backprojection_adjoint",
    "siddon_maplors", "siddon_osem_subset", "siddon_count_nnz", "mlem_update",
]


def _stream():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(x):
    return int(x) if x else None


def _check_dev(t, name):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be on a GPU (got {t.device}); srf_amd has no CPU path")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32 (got {t.dtype})")


def _soa(lors, ncols):
    """Return (tensor, n, ld) for a [>=ncols, N] SoA LOR tensor."""
    _check_dev(lors, "lors")
    if lors.dim() != 2 or lors.shape[0] < ncols:
        raise ValueError(f"lors must be [{ncols}, N] (got {tuple(lors.shape)})")
    if lors.stride(1) != 1 or lors.stride(0) < lors.shape[1]:
        lors = lors.contiguous()
    return lors, lors.shape[1], max(lors.stride(0), lors.shape[1])


def _vec(v, n, name):
    _check_dev(v, name)
    v = v.reshape(-1)
    if v.numel() != n:
        raise ValueError(f"{name} has {v.numel()} values for {n} LORs")
    return v.contiguous()


def _geom(grid, center, size, strides=(0, 0, 0)):
    grid = [int(g) for g in _host(grid)]
    center = [float(c) for c in _host(center)]
    size = [float(s) for s in _host(size)]
    if len(grid) != 3 or len(center) != 3 or len(size) != 3:
        raise ValueError("grid, center and size need 3 entries")
    return grid, _lib.make_grid(grid, center, size, strides)


def _host(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().tolist()
    return list(x)


def _image_like(image, grid):
    _check_dev(image, "image")
    n = grid[0] * grid[1] * grid[2]
    if image.numel() != n:
        raise ValueError(f"image has {image.numel()} voxels, grid {grid} needs {n}")
    return image.contiguous()


def _out_image(out, grid, like):
    """A caller-supplied output image: a float32 device tensor on the input's
    device, contiguous, with exactly the grid's voxel count (the kernels write
    gx * gy * gz floats from its first element)."""
    _check_dev(out, "out")
    n = grid[0] * grid[1] * grid[2]
    if out.device != like.device:
        raise ValueError(f"out is on {out.device}, the inputs on {like.device}")
    if not out.is_contiguous():
        raise ValueError("out must be contiguous")
    if out.numel() != n:
        raise ValueError(f"out has {out.numel()} voxels, grid {grid} needs {n}")
    return out


def _ptr(t):
    return ctypes_void(t.data_ptr()) if t is not None else None


# ------------------------------------------------------------------- TOR ---

def projection(lors, image, grid, center, size, kernel_width, geo_sigma2_flag=False,
               tof_bin=1.0e4, tof_sigma2=1.0e4, full_patch=False, schedule=None):
    """Op ``Projection`` of tor.so (tor.cc:14-28, 106-181).  full_patch walks
    every voxel of the reference's patch (verification path); schedule is
    (lanes_per_event, block) or None for the default."""
    lors, n, ld = _soa(lors, 10)
    g, cg = _geom(grid, center, size)
    image = _image_like(image, g)
    out = torch.empty(n, dtype=torch.float32, device=lors.device)
    m = _lib.make_tor_model(kernel_width, geo_sigma2_flag, tof_bin, tof_sigma2)
    _lib.call("srfx_tor_project", _ptr(lors), n, ld, _ptr(image), cg, m, _ptr(out),
              (_lib.SRFX_FULL_PATCH if full_patch else 0) | _lib.schedule_flags(schedule),
              _stream())
    return out


def backprojection(image, lors, lors_value, grid, center, size, tof_bin=1.0e4, tof_sigma2=1.0e4,
                   kernel_width=None, geo_sigma2_flag=False, full_patch=False, schedule=None):
    """Op ``Backprojection`` of tor.so (tor.cc:30-47, 203-271): returns an
    image shaped like ``image`` holding sum_i w_ij / v_i over v_i > 1e-7."""
    if kernel_width is None:
        raise ValueError("kernel_width is required")
    lors, n, ld = _soa(lors, 10)
    g, cg = _geom(grid, center, size)
    _image_like(image, g)
    v = _vec(lors_value, n, "lors_value")
    out = torch.empty_like(image, memory_format=torch.contiguous_format)
    m = _lib.make_tor_model(kernel_width, geo_sigma2_flag, tof_bin, tof_sigma2)
    _lib.call("srfx_tor_backproject", _ptr(lors), n, ld, _ptr(v), cg, m, _ptr(out),
              (_lib.SRFX_FULL_PATCH if full_patch else 0) | _lib.schedule_flags(schedule),
              _stream())
    return out


def backprojection_adjoint(image, lors, g_values, grid, center, size, tof_bin=1.0e4,
                           tof_sigma2=1.0e4, kernel_width=None, geo_sigma2_flag=False,
                           schedule=None):
    """Plain adjoint sum_i w_ij g_i (no reference counterpart)."""
    if kernel_width is None:
        raise ValueError("kernel_width is required")
    lors, n, ld = _soa(lors, 10)
    g, cg = _geom(grid, center, size)
    _image_like(image, g)
    v = _vec(g_values, n, "g")
    out = torch.empty_like(image, memory_format=torch.contiguous_format)
    m = _lib.make_tor_model(kernel_width, geo_sigma2_flag, tof_bin, tof_sigma2)
    _lib.call("srfx_tor_backproject_adjoint", _ptr(lors), n, ld, _ptr(v), cg, m, _ptr(out),
              _lib.schedule_flags(schedule), _stream())
    return out


def maplors(image, grid, center, size, lors, lors_value, kernel_width, geo_sigma2_flag=False,
            schedule=None):
    """Op ``Maplors`` of tor.so (tor.cc:49-64, 290-348): lors [7, N] =
    x1..z2, sigma2 factor; sum_i w_ij / v_i without TOF and threshold."""
    lors, n, ld = _soa(lors, 7)
    g, cg = _geom(grid, center, size)
    _image_like(image, g)
    v = _vec(lors_value, n, "lors_value")
    out = torch.empty_like(image, memory_format=torch.contiguous_format)
    _lib.call("srfx_tor_maplors", _ptr(lors), n, ld, _ptr(v), cg, float(kernel_width),
              1 if geo_sigma2_flag else 0, _ptr(out), _lib.schedule_flags(schedule), _stream())
    return out


def osem_subset(lors, image, grid, center, size, kernel_width, geo_sigma2_flag=False,
                tof_bin=1.0e4, tof_sigma2=1.0e4, out=None, accumulate=False, schedule=None):
    """Fused projection + ratio backprojection of one group (no reference op:
    equals backprojection(lors_value=projection(...)))."""
    lors, n, ld = _soa(lors, 10)
    g, cg = _geom(grid, center, size)
    image = _image_like(image, g)
    if out is None:
        out = torch.empty_like(image, memory_format=torch.contiguous_format)
        accumulate = False
    else:
        out = _out_image(out, g, image)
    m = _lib.make_tor_model(kernel_width, geo_sigma2_flag, tof_bin, tof_sigma2)
    _lib.call("srfx_tor_osem_subset", _ptr(lors), n, ld, _ptr(image), cg, cg, m, _ptr(out),
              (_lib.SRFX_ACCUMULATE if accumulate else 0) | _lib.schedule_flags(schedule),
              _stream())
    return out


def count_nnz(lors, grid, center, size, kernel_width, geo_sigma2_flag=False):
    """Per-event count of non-zero TOR system-matrix entries (int32 [N]) and
    their total (int64 scalar tensor)."""
    lors, n, ld = _soa(lors, 10)
    g, cg = _geom(grid, center, size)
    per = torch.empty(n, dtype=torch.int32, device=lors.device)
    tot = torch.zeros(1, dtype=torch.int64, device=lors.device)
    m = _lib.make_tor_model(kernel_width, geo_sigma2_flag, 0.0, 1.0e6)
    _lib.call("srfx_tor_count_nnz", _ptr(lors), n, ld, cg, m, _ptr(per), _ptr(tot), _stream())
    return per, tot


# ---------------------------------------------------------------- Siddon ---

def siddon_projection(lors, image, grid, center, size, tof_bin=1.0e4, tof_sigma2=1.0e4):
    """Op ``Projection`` of siddon.so (siddon.cc:14-26): lors [7, N] = x1..z2, tof."""
    lors, n, ld = _soa(lors, 7)
    g, cg = _geom(grid, center, size)
    image = _image_like(image, g)
    out = torch.empty(n, dtype=torch.float32, device=lors.device)
    _lib.call("srfx_siddon_project", _ptr(lors), n, ld, _ptr(image), cg, float(tof_bin),
              float(tof_sigma2), _ptr(out), 0, _stream())
    return out


def siddon_backprojection(image, lors, lors_value, grid, center, size, tof_bin=1.0e4,
                          tof_sigma2=1.0e4):
    """Op ``Backprojection`` of siddon.so (siddon.cc:28-43): sum a_ij / v_i, v_i > 0."""
    lors, n, ld = _soa(lors, 7)
    g, cg = _geom(grid, center, size)
    _image_like(image, g)
    v = _vec(lors_value, n, "lors_value")
    out = torch.empty_like(image, memory_format=torch.contiguous_format)
    _lib.call("srfx_siddon_backproject", _ptr(lors), n, ld, _ptr(v), cg, float(tof_bin),
              float(tof_sigma2), _ptr(out), 0, _stream())
    return out


def siddon_backprojection_adjoint(image, lors, g_values, grid, center, size, tof_bin=1.0e4,
                                  tof_sigma2=1.0e4):
    lors, n, ld = _soa(lors, 7)
    g, cg = _geom(grid, center, size)
    _image_like(image, g)
    v = _vec(g_values, n, "g")
    out = torch.empty_like(image, memory_format=torch.contiguous_format)
    _lib.call("srfx_siddon_backproject_adjoint", _ptr(lors), n, ld, _ptr(v), cg,
              float(tof_bin), float(tof_sigma2), _ptr(out), 0, _stream())
    return out


def siddon_maplors(image, grid, center, size, lors, lors_value):
    """Op ``Maplors`` of siddon.so (siddon.cc:45-57): non-TOF, 1/(L*L) weight."""
    lors, n, ld = _soa(lors, 6)
    g, cg = _geom(grid, center, size)
    _image_like(image, g)
    v = _vec(lors_value, n, "lors_value")
    out = torch.empty_like(image, memory_format=torch.contiguous_format)
    _lib.call("srfx_siddon_maplors", _ptr(lors), n, ld, _ptr(v), cg, _ptr(out), 0, _stream())
    return out


def siddon_osem_subset(lors, image, grid, center, size, tof_bin=1.0e4, tof_sigma2=1.0e4,
                       out=None, accumulate=False):
    lors, n, ld = _soa(lors, 7)
    g, cg = _geom(grid, center, size)
    image = _image_like(image, g)
    if out is None:
        out = torch.empty_like(image, memory_format=torch.contiguous_format)
        accumulate = False
    else:
        out = _out_image(out, g, image)
    _lib.call("srfx_siddon_osem_subset", _ptr(lors), n, ld, _ptr(image), cg, float(tof_bin),
              float(tof_sigma2), _ptr(out), _lib.SRFX_ACCUMULATE if accumulate else 0, _stream())
    return out


def siddon_count_nnz(lors, grid, center, size, tof_bin=1.0e4, tof_sigma2=1.0e4):
    lors, n, ld = _soa(lors, 7)
    g, cg = _geom(grid, center, size)
    per = torch.empty(n, dtype=torch.int32, device=lors.device)
    tot = torch.zeros(1, dtype=torch.int64, device=lors.device)
    _lib.call("srfx_siddon_count_nnz", _ptr(lors), n, ld, cg, float(tof_bin), float(tof_sigma2),
              _ptr(per), _ptr(tot), _stream())
    return per, tot


def mlem_update(x, emap, bp):
    """x <- x * emap * bp in place (model/recon_step.py:48-49)."""
    for t, nm in ((x, "x"), (emap, "emap"), (bp, "bp")):
        _check_dev(t, nm)
        if not t.is_contiguous():
            raise ValueError(f"{nm} must be contiguous")
    if not (x.numel() == emap.numel() == bp.numel()):
        raise ValueError("x, emap and bp must have the same number of voxels")
    _lib.call("srfx_mlem_update", _ptr(x), _ptr(emap), _ptr(bp), x.numel(), _stream())
    return x
