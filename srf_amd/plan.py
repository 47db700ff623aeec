"""Tiled TOR ops over a fixed event set (srfx_tor_plan_*).

A ``TorPlan`` bins the events of one group frame into image tiles of
``slices`` x ``wx`` x ``wy`` op-frame voxels once; its ops then compute the
same projection / backprojection / Maplors / OSEM step as the untiled ops of
``srf_amd.ops`` (src/cpp/tf_custom_ops/tor.cc:14-64 semantics) with the
scatter accumulated per tile in LDS and flushed with coalesced atomics, and
the projection optionally gathered from tiles staged in LDS.  OSEM reuses one
plan per (subset, group) for every iteration, since the events of a subset
do not change between iterations (graph/reconstruction/worker.py:79-85).

The plan keeps a reference to the LOR tensor: it must not be modified while
the plan lives.  Images use the plan's grid layout: contiguous [gz][gy][gx] in
the op frame, or the element strides given at construction.

``SiddonPlan`` (srfx_siddon_plan_*) does the same for the Siddon model's
backprojection: each LOR's walk is cut into per-tile runs once, and every
backprojection accumulates the runs of a tile in LDS.
"""
import ctypes

import torch

from . import _lib
from .ops import _check_dev, _geom, _ptr, _soa, _stream, _vec

DEFAULT_TILE = (12, 32, 10)
# Chord walks (events whose candidate box would be wider than 4 voxels) list
# fewer (event, tile) pairs in deeper windows, and have no single-walk kernel
# to fit five workgroups per CU.
CHORD_TILE = (12, 32, 12)


def default_tile(grid, size, kernel_width):
    """The default tile of a plan over a frame of ``grid`` voxels spanning
    ``size`` mm at ``kernel_width``: DEFAULT_TILE for the candidate-box
    walks, CHORD_TILE for the chord walk."""
    return DEFAULT_TILE if _lib.tor_patch_info(grid, size, kernel_width)[1] else CHORD_TILE
DEFAULT_CHUNK = 2048


class TorPlan:
    def __init__(self, lors, grid, center, size, kernel_width, geo_sigma2_flag=False,
                 tof_bin=1.0e4, tof_sigma2=1.0e4, strides=(0, 0, 0), ncols=10,
                 tile=DEFAULT_TILE, chunk=DEFAULT_CHUNK, block=256, axis=None,
                 tiled_projection=False):
        if ncols not in (7, 10):
            raise ValueError("ncols must be 10 (TOF rows) or 7 (Maplors rows)")
        self.lors, self.n, ld = _soa(lors, ncols)
        self.grid, self._cgrid = _geom(grid, center, size, strides)
        self.center, self.size = [float(c) for c in center], [float(s) for s in size]
        self.strides = tuple(int(s) for s in strides)
        self.axis = axis
        self.ncols = ncols
        self.tile = tuple(int(t) for t in tile)
        model = _lib.make_tor_model(kernel_width, geo_sigma2_flag, tof_bin, tof_sigma2)
        self.tiled_projection = bool(tiled_projection)
        t = _lib.SrfxTile(self.tile[0], self.tile[1], self.tile[2], int(chunk), int(block),
                          1 if tiled_projection else 0)
        h = ctypes.c_void_p()
        _lib.call("srfx_tor_plan_create", _ptr(self.lors), self.n, ld, ncols, self._cgrid, model,
                  t, ctypes.byref(h), _stream())
        self._h = h

    @classmethod
    def subsets(cls, lors, bounds, grid, center, size, kernel_width, geo_sigma2_flag=False,
                tof_bin=1.0e4, tof_sigma2=1.0e4, strides=(0, 0, 0), ncols=10,
                tile=DEFAULT_TILE, chunk=DEFAULT_CHUNK, block=256, axis=None):
        """Plans of the event ranges bounds[k]:bounds[k + 1] of one LOR tensor
        (the OSEM subsets of one axis group), built together by
        srfx_tor_plan_create_subsets.  Returns one TorPlan per range."""
        if ncols not in (7, 10):
            raise ValueError("ncols must be 10 (TOF rows) or 7 (Maplors rows)")
        full, n, ld = _soa(lors, ncols)
        bounds = [int(b) for b in bounds]
        K = len(bounds) - 1
        if K < 1 or bounds[0] < 0 or bounds[-1] > n or any(
                b1 < b0 for b0, b1 in zip(bounds, bounds[1:])):
            raise ValueError("bounds must be non-decreasing event offsets within the LORs")
        g, cgrid = _geom(grid, center, size, strides)
        model = _lib.make_tor_model(kernel_width, geo_sigma2_flag, tof_bin, tof_sigma2)
        t = _lib.SrfxTile(*[int(v) for v in tile], int(chunk), int(block), 0)
        hs = (ctypes.c_void_p * K)()
        bs = (ctypes.c_int64 * (K + 1))(*bounds)
        _lib.call("srfx_tor_plan_create_subsets", _ptr(full), n, ld, ncols, cgrid, model, t, K,
                  bs, hs, _stream())
        return [cls._wrap(hs[k], full, bounds[k], bounds[k + 1], g, cgrid, center, size,
                          strides, axis, ncols, tile) for k in range(K)]

    @classmethod
    def subset_groups(cls, groups, kernel_width, geo_sigma2_flag=False, tof_bin=1.0e4,
                      tof_sigma2=1.0e4, ncols=10, chunk=DEFAULT_CHUNK, block=256):
        """``subsets`` for several LOR tensors in one build
        (srfx_tor_plan_create_subset_groups): ``groups`` is a list of dicts
        with the keys lors, bounds, grid, center, size and optionally strides,
        tile and axis, as ``subsets`` takes them.  Returns, per group, its list
        of TorPlan."""
        if ncols not in (7, 10):
            raise ValueError("ncols must be 10 (TOF rows) or 7 (Maplors rows)")
        model = _lib.make_tor_model(kernel_width, geo_sigma2_flag, tof_bin, tof_sigma2)
        specs, keep = [], []
        arr = (_lib.SrfxSubsetGroup * max(len(groups), 1))()
        for i, gr in enumerate(groups):
            full, n, ld = _soa(gr["lors"], ncols)
            bounds = [int(b) for b in gr["bounds"]]
            K = len(bounds) - 1
            if K < 1 or bounds[0] < 0 or bounds[-1] > n or any(
                    b1 < b0 for b0, b1 in zip(bounds, bounds[1:])):
                raise ValueError("bounds must be non-decreasing event offsets within the LORs")
            strides = gr.get("strides", (0, 0, 0))
            tile = gr.get("tile", DEFAULT_TILE)
            g, cgrid = _geom(gr["grid"], gr["center"], gr["size"], strides)
            bs = (ctypes.c_int64 * (K + 1))(*bounds)
            keep.append(bs)
            arr[i] = _lib.SrfxSubsetGroup(
                _ptr(full), n, ld, cgrid,
                _lib.SrfxTile(*[int(v) for v in tile], int(chunk), int(block), 0), K, bs)
            specs.append((full, bounds, g, cgrid, gr["center"], gr["size"], strides,
                          gr.get("axis"), tile))
        total = sum(len(sp[1]) - 1 for sp in specs)
        hs = (ctypes.c_void_p * max(total, 1))()
        _lib.call("srfx_tor_plan_create_subset_groups", arr, len(groups), ncols, model, hs,
                  _stream())
        out, j = [], 0
        for full, bounds, g, cgrid, center, size, strides, axis, tile in specs:
            plans = []
            for k in range(len(bounds) - 1):
                plans.append(cls._wrap(hs[j], full, bounds[k], bounds[k + 1], g, cgrid, center,
                                       size, strides, axis, ncols, tile))
                j += 1
            out.append(plans)
        return out

    @classmethod
    def _wrap(cls, handle, full, b0, b1, g, cgrid, center, size, strides, axis, ncols, tile):
        """The TorPlan of a handle from a subset build (events b0:b1 of full)."""
        pl = cls.__new__(cls)
        pl.lors = full[:, b0:b1]
        pl.n = b1 - b0
        pl.grid, pl._cgrid = g, cgrid
        pl.center, pl.size = [float(c) for c in center], [float(s) for s in size]
        pl.strides = tuple(int(s) for s in strides)
        pl.axis = axis
        pl.ncols = ncols
        pl.tile = tuple(int(v) for v in tile)
        pl.tiled_projection = False
        pl._h = ctypes.c_void_p(handle)
        return pl

    def info(self):
        """(number of (event, tile) listings, number of workgroup chunks)."""
        a, b = ctypes.c_int64(), ctypes.c_int32()
        _lib.call("srfx_tor_plan_info", self._h, ctypes.byref(a), ctypes.byref(b))
        return int(a.value), int(b.value)

    def walks(self):
        """(candidate box 3 or 4, or 0 for chord walks; whether every event
        walks the NB x (NB - 1) box, so that the scatter can run its
        single-walk kernel)."""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _lib.call("srfx_tor_plan_walks", self._h, ctypes.byref(a), ctypes.byref(b))
        return int(a.value), bool(b.value)

    def device_bytes(self):
        """Device memory held by the plan: 4 bytes per listing (its event),
        148 per event (144-byte walk record, projection buffer), 16 per chunk,
        and 12 more per listing and 4 per event with the tiled-projection
        tables."""
        nl, nc = self.info()
        return 4 * nl + 148 * self.n + 16 * nc + (12 * nl + 4 * self.n if self.tiled_projection
                                                  else 0)

    def _image(self, image, name="image"):
        _check_dev(image, name)
        g = self.grid
        if self.strides == (0, 0, 0):
            if image.numel() != g[0] * g[1] * g[2] or not image.is_contiguous():
                raise ValueError(f"{name} must be a contiguous {g[::-1]} tensor")
        return image

    def _out(self, out, like):
        if out is None:
            if self.strides != (0, 0, 0):
                raise ValueError("strided plans need an explicit output image")
            g = self.grid
            return torch.empty((g[2], g[1], g[0]), dtype=torch.float32, device=like.device)
        return self._image(out, "out")

    def projection(self, image, op_frame=False, schedule=None):
        """Projection of the plan's events: the tiled gather for plans built
        with tiled_projection, else the walk from the event records (the OSEM
        step's projection kernel).  op_frame: image is the contiguous op-frame
        copy [gz][gy][gx] (frame gathers)."""
        if op_frame:
            _check_dev(image, "image")
            g = self.grid
            if tuple(image.shape) != (g[2], g[1], g[0]) or not image.is_contiguous():
                raise ValueError(f"op-frame image must be a contiguous {g[::-1]} tensor")
        else:
            image = self._image(image)
        out = torch.empty(self.n, dtype=torch.float32, device=image.device)
        flags = _lib.schedule_flags(schedule) | (_lib.SRFX_IMAGE_OP_FRAME if op_frame else 0)
        _lib.call("srfx_tor_plan_project", self._h, _ptr(image), _ptr(out), flags, _stream())
        return out

    def _scatter(self, name, values, out, accumulate, device):
        v = _vec(values, self.n, "lors_value")
        out = self._out(out, v)
        flags = _lib.SRFX_ACCUMULATE if accumulate else 0
        _lib.call(name, self._h, _ptr(v), _ptr(out), flags, _stream())
        return out

    def backprojection(self, lors_value, out=None, accumulate=False):
        return self._scatter("srfx_tor_plan_backproject", lors_value, out, accumulate, None)

    def backprojection_adjoint(self, g, out=None, accumulate=False):
        return self._scatter("srfx_tor_plan_backproject_adjoint", g, out, accumulate, None)

    def maplors(self, lors_value, out=None, accumulate=False):
        return self._scatter("srfx_tor_plan_maplors", lors_value, out, accumulate, None)

    def osem_subset(self, image, out=None, accumulate=False, tiled_projection=False,
                    schedule=None, out_strides=None, stage=None, op_frame=False):
        """Backprojection of 1/p (p > 1e-7) with p the projection of image.
        stage "project" / "scatter" runs only that half (the projection is
        kept in the plan's event records in between).  op_frame: image is the
        contiguous op-frame copy [gz][gy][gx] (any plan strides); out keeps
        the plan's strides (or out_strides)."""
        if op_frame:
            _check_dev(image, "image")
            g = self.grid
            if tuple(image.shape) != (g[2], g[1], g[0]) or not image.is_contiguous():
                raise ValueError(f"op-frame image must be a contiguous {g[::-1]} tensor")
            if out is None:
                raise ValueError("op_frame needs an explicit output image")
        else:
            image = self._image(image)
        out = self._out(out, image)
        og = self._cgrid if out_strides is None else _lib.make_grid(
            self.grid, self.center, self.size, out_strides)
        st = {None: 0, "project": _lib.SRFX_OSEM_PROJECT_ONLY,
              "scatter": _lib.SRFX_OSEM_SCATTER_ONLY}[stage]
        flags = ((_lib.SRFX_ACCUMULATE if accumulate else 0)
                 | (_lib.SRFX_TILED_PROJECTION if tiled_projection else 0)
                 | _lib.schedule_flags(schedule) | st
                 | (_lib.SRFX_IMAGE_OP_FRAME if op_frame else 0))
        _lib.call("srfx_tor_plan_osem_subset", self._h, _ptr(image), og, _ptr(out), flags,
                  _stream())
        return out

    def scales(self):
        """The backprojection scales the last OSEM step stored in the event
        records: 1/p for p > 1e-7, else 0 (float32 [n])."""
        out = torch.empty(self.n, dtype=torch.float32,
                          device=self.lors.device)
        _lib.call("srfx_tor_plan_scales", self._h, _ptr(out), _stream())
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.call("srfx_tor_plan_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


DEFAULT_SIDDON_TILE = (16, 16, 16)
DEFAULT_SIDDON_CHUNK = 2048


class SiddonPlan:
    """Tiled Siddon backprojection over a fixed LOR set (srfx_siddon_plan_*).

    The plan cuts each LOR's walk (siddon.cu.cc:293-331) into runs, one per
    image tile of ``tile`` voxels it crosses, and lists the runs by tile; its
    backprojections accumulate each tile in LDS and flush it with coalesced
    atomics.  Same voxels and segment values as the untiled ops
    (siddon.cc:28-43 semantics), summed in another order.  OSEM reuses one
    plan per subset for every iteration.

    The plan keeps a reference to the LOR tensor (7 x N SoA: x1..z2, tof): it
    must not be modified while the plan lives.  Images are (nx, ny, nz)
    world-frame tensors with the plan's element strides.
    """

    def __init__(self, lors, grid, center, size, tof_bin=1.0e4, tof_sigma2=1.0e4,
                 strides=None, tile=DEFAULT_SIDDON_TILE, chunk=DEFAULT_SIDDON_CHUNK, block=256,
                 sort=True):
        self.lors, self.n, ld = _soa(lors, 7)
        self._layout(grid, center, size, strides, tile)
        self.sorted = bool(sort)
        t = _lib.SrfxSiddonTile(*self.tile, int(chunk), int(block), int(self.sorted))
        h = ctypes.c_void_p()
        _lib.call("srfx_siddon_plan_create", _ptr(self.lors), self.n, ld, self._cgrid,
                  float(tof_bin), float(tof_sigma2), t, ctypes.byref(h), _stream())
        self._h = h

    @classmethod
    def subsets(cls, lors, bounds, grid, center, size, tof_bin=1.0e4, tof_sigma2=1.0e4,
                strides=None, tile=DEFAULT_SIDDON_TILE, chunk=DEFAULT_SIDDON_CHUNK, block=256,
                sort=True):
        """Plans of the LOR ranges bounds[k]:bounds[k + 1] of one LOR tensor
        (the OSEM subsets), built together by srfx_siddon_plan_create_subsets.
        Returns one SiddonPlan per range."""
        full, n, ld = _soa(lors, 7)
        bounds = [int(b) for b in bounds]
        K = len(bounds) - 1
        if K < 1 or bounds[0] < 0 or bounds[-1] > n or any(
                b1 < b0 for b0, b1 in zip(bounds, bounds[1:])):
            raise ValueError("bounds must be non-decreasing LOR offsets within the LORs")
        proto = cls.__new__(cls)
        proto._layout(grid, center, size, strides, tile)
        proto.sorted = bool(sort)
        t = _lib.SrfxSiddonTile(*proto.tile, int(chunk), int(block), int(proto.sorted))
        hs = (ctypes.c_void_p * K)()
        bs = (ctypes.c_int64 * (K + 1))(*bounds)
        _lib.call("srfx_siddon_plan_create_subsets", _ptr(full), n, ld, proto._cgrid,
                  float(tof_bin), float(tof_sigma2), t, K, bs, hs, _stream())
        out = []
        for k in range(K):
            pl = cls.__new__(cls)
            pl.__dict__.update({a: getattr(proto, a) for a in
                                ("strides", "grid", "_cgrid", "_dense", "tile", "sorted")})
            pl.lors = full[:, bounds[k]:bounds[k + 1]]
            pl.n = bounds[k + 1] - bounds[k]
            pl._h = ctypes.c_void_p(hs[k])
            out.append(pl)
        return out

    def _layout(self, grid, center, size, strides, tile):
        grid = [int(g) for g in grid]
        if strides is None:
            strides = (grid[1] * grid[2], grid[2], 1)
        self.strides = tuple(int(s) for s in strides)
        self.grid, self._cgrid = _geom(grid, center, size, self.strides)
        # strides of a dense layout (some axis order of a contiguous array)
        order = sorted(range(3), key=lambda k: self.strides[k])
        dense, step = True, 1
        for k in order:
            dense &= self.strides[k] == step
            step *= grid[k]
        self._dense = dense
        self.tile = tuple(int(t) for t in tile)

    def info(self):
        """(number of (LOR, tile) runs, number of workgroup chunks)."""
        a, b = ctypes.c_int64(), ctypes.c_int32()
        _lib.call("srfx_siddon_plan_info", self._h, ctypes.byref(a), ctypes.byref(b))
        return int(a.value), int(b.value)

    def device_bytes(self):
        """Device memory held by the plan: 48 bytes per run, 52 per LOR (ray
        record, projection buffer), 16 per chunk, 8 per tile, and with the
        sorted LOR stream 32 more per LOR (sorted copy, permutation) and the
        build's sort buffers (24 per LOR)."""
        nr, nc = self.info()
        nt = 1
        for g, t in zip(self.grid, self.tile):
            nt *= -(-g // min(t, g))
        return (48 * nr + 52 * self.n + 16 * nc + 8 * nt
                + (56 * self.n if self.sorted else 0))

    def _image(self, image, name):
        """image: a (nx, ny, nz) tensor with the plan's strides, or, for a
        plan whose strides describe a dense layout, any contiguous tensor of
        that layout (e.g. [gz][gy][gx] for strides (1, gx, gx gy))."""
        _check_dev(image, name)
        g = self.grid
        if list(image.shape) == g and tuple(image.stride()) == self.strides:
            return image
        if self._dense and image.is_contiguous() and image.numel() == g[0] * g[1] * g[2]:
            return image
        raise ValueError(f"{name} must be a {tuple(g)} tensor with strides {self.strides}")

    def _out(self, out, like):
        if out is None:
            out = torch.empty_strided(self.grid, self.strides, dtype=torch.float32,
                                      device=like.device)
        return self._image(out, "out")

    def backprojection(self, lors_value, out=None, accumulate=False):
        v = _vec(lors_value, self.n, "lors_value")
        out = self._out(out, v)
        _lib.call("srfx_siddon_plan_backproject", self._h, _ptr(v), _ptr(out),
                  _lib.SRFX_ACCUMULATE if accumulate else 0, _stream())
        return out

    def backprojection_adjoint(self, g, out=None, accumulate=False):
        v = _vec(g, self.n, "g")
        out = self._out(out, v)
        _lib.call("srfx_siddon_plan_backproject_adjoint", self._h, _ptr(v), _ptr(out),
                  _lib.SRFX_ACCUMULATE if accumulate else 0, _stream())
        return out

    def osem_subset(self, image, out=None, accumulate=False, counts=None, stage=None):
        """Backprojection of 1/p (p > 0), or of c/p with per-LOR counts c
        (float32 [n], LOR order), with p the projection of image.  stage
        "project" / "scatter" runs only that half (the projection stays in the
        plan in between)."""
        flags = ((_lib.SRFX_ACCUMULATE if accumulate else 0)
                 | {None: 0, "project": _lib.SRFX_OSEM_PROJECT_ONLY,
                    "scatter": _lib.SRFX_OSEM_SCATTER_ONLY}[stage])
        c = None if counts is None else _vec(counts, self.n, "counts")
        if stage == "project":
            image = self._image(image, "image")
            _lib.call("srfx_siddon_plan_osem_subset", self._h, _ptr(image), None, None, flags,
                      _stream())
            return None
        if stage != "scatter":
            image = self._image(image, "image")
        out = self._out(out, c if c is not None else self.lors)
        _lib.call("srfx_siddon_plan_osem_subset", self._h,
                  None if stage == "scatter" else _ptr(image),
                  None if c is None else _ptr(c), _ptr(out), flags, _stream())
        return out

    def projection(self):
        """The projection of the last osem_subset call (float32 [n])."""
        out = torch.empty(self.n, dtype=torch.float32, device=self.lors.device)
        _lib.call("srfx_siddon_plan_projection", self._h, _ptr(out), _stream())
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.call("srfx_siddon_plan_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
