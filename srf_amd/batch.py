"""Batched TOF OSEM: independent reconstructions advanced together.

``BatchedOSEM(recons, model, nb_subsets)`` plans a batch of reconstructions
(each a dict with ``image`` (Image), ``emap`` (Image of the inverse efficiency
map) and ``data`` (ListModeDataSplit from preprocess.recon_process)) that share
one SplitLoRsModel.  ``run_subset(k)`` advances every reconstruction by OSEM
subset k (srfx_tor_batch_subset: four kernels, two of them persistent);
``capture()`` records one full iteration (all subsets) into a HIP graph that
``replay()`` reruns.  Every (reconstruction, axis group, subset) gets a tile
plan at creation; a subset step is a persistent projection over the events of
all reconstructions, a persistent tiled backprojection over the tile chunks of
all plans and the batched update.  ``serial_iterate`` runs the same
reconstructions one after another through the per-reconstruction OSEM driver
(projection + tiled backprojection), for comparison.
"""
import ctypes

import torch

from . import _lib
from .ops import _check_dev, _stream
from .recon import OSEM


# Schedules of the persistent projection: lanes per event x events per work item.
LANES = (1, 4, 16)
ITEM_EVENTS = (64, 128, 256)


class SrfxRecon(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int32 * 3), ("center", ctypes.c_float * 3),
                ("size", ctypes.c_float * 3), ("image", ctypes.c_void_p),
                ("emap", ctypes.c_void_p), ("bp", ctypes.c_void_p),
                ("lors", ctypes.c_void_p * 3), ("n", ctypes.c_int64 * 3),
                ("ld", ctypes.c_int64 * 3)]


def _bind():
    lib = _lib.load()
    if not getattr(lib, "_batch_bound", False):
        lib.srfx_tor_batch_create.argtypes = [ctypes.POINTER(SrfxRecon), ctypes.c_int32,
                                              ctypes.c_int32, _lib.SrfxTorModel,
                                              ctypes.POINTER(ctypes.c_void_p)]
        lib.srfx_tor_batch_create.restype = ctypes.c_int
        lib.srfx_tor_batch_subset.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
        lib.srfx_tor_batch_subset.restype = ctypes.c_int
        lib.srfx_tor_batch_set_schedule.argtypes = [ctypes.c_void_p, ctypes.c_int32,
                                                    ctypes.c_int32]
        lib.srfx_tor_batch_set_schedule.restype = ctypes.c_int
        lib.srfx_tor_batch_destroy.argtypes = [ctypes.c_void_p]
        lib.srfx_tor_batch_destroy.restype = ctypes.c_int
        lib._batch_bound = True
    return lib


def _check(lib, name, status):
    if status != _lib.SRFX_OK:
        raise _lib.SrfxError(name, status, lib.srfx_strerror(status).decode())


class BatchedOSEM:
    def __init__(self, recons, model, nb_subsets):
        self.lib = _bind()
        self.recons = recons
        self.model = model
        self.nb_subsets = int(nb_subsets)
        self.bps = []
        arr = (SrfxRecon * len(recons))()
        for r, rc in zip(arr, recons):
            x, e = rc["image"].data, rc["emap"].data
            for t, nm in ((x, "image"), (e, "emap")):
                _check_dev(t, nm)
                if not t.is_contiguous() or t.dim() != 3:
                    raise ValueError(f"{nm} must be a contiguous (nx, ny, nz) tensor")
            if e.shape != x.shape:
                raise ValueError("emap and image shapes differ")
            bp = torch.empty_like(x)
            self.bps.append(bp)
            for k in range(3):
                r.shape[k] = x.shape[k]
                r.center[k] = rc["image"].center[k]
                r.size[k] = rc["image"].size[k]
            r.image, r.emap, r.bp = x.data_ptr(), e.data_ptr(), bp.data_ptr()
            for k, a in enumerate("xyz"):
                soa = rc["data"][a].soa
                _check_dev(soa, "lors")
                if soa.dim() != 2 or soa.shape[0] < 10 or (soa.shape[1] and soa.stride(1) != 1):
                    raise ValueError("group LORs must be [10, N] row-contiguous")
                r.lors[k] = soa.data_ptr() if soa.shape[1] else None
                r.n[k] = soa.shape[1]
                r.ld[k] = max(soa.stride(0), soa.shape[1])
        self._arr = arr
        handle = ctypes.c_void_p()
        _check(self.lib, "srfx_tor_batch_create",
               self.lib.srfx_tor_batch_create(arr, len(recons), self.nb_subsets, model._model(),
                                              ctypes.byref(handle)))
        self.handle = handle
        self.graph = None
        self.schedule = (4, 64)

    def set_schedule(self, lanes_per_event, events_per_item):
        """Schedule of the persistent projection (srfx_tor_batch_set_schedule):
        lanes per event in LANES, events per work item in ITEM_EVENTS.  A
        graph keeps the kernel and grid it recorded, so a captured graph is
        dropped: replay() then raises until capture() records the new
        schedule."""
        _check(self.lib, "srfx_tor_batch_set_schedule",
               self.lib.srfx_tor_batch_set_schedule(self.handle, int(lanes_per_event),
                                                    int(events_per_item)))
        self.schedule = (int(lanes_per_event), int(events_per_item))
        self.graph = None

    def autotune(self, reps=2):
        """Time one OSEM iteration (all subsets) for every schedule in
        LANES x ITEM_EVENTS, keep the fastest and restore every image to its
        value before the call.  Returns {(lanes, events): ms}, the minimum of
        ``reps`` timed iterations after one warm-up each."""
        saved = [rc["image"].data.clone() for rc in self.recons]

        def restore():
            for rc, x in zip(self.recons, saved):
                rc["image"].data.copy_(x)

        table = {}
        for lanes in LANES:
            for events in ITEM_EVENTS:
                self.set_schedule(lanes, events)
                best = float("inf")
                for rep in range(reps + 1):
                    restore()
                    a = torch.cuda.Event(enable_timing=True)
                    b = torch.cuda.Event(enable_timing=True)
                    a.record()
                    self.iterate(1)
                    b.record()
                    torch.cuda.synchronize()
                    if rep:
                        best = min(best, a.elapsed_time(b))
                table[(lanes, events)] = best
        restore()
        self.set_schedule(*min(table, key=table.get))
        torch.cuda.synchronize()
        return table

    def run_subset(self, k):
        _check(self.lib, "srfx_tor_batch_subset",
               self.lib.srfx_tor_batch_subset(self.handle, int(k), _stream()))

    def iterate(self, nb_iterations=1):
        for _ in range(nb_iterations):
            for k in range(self.nb_subsets):
                self.run_subset(k)

    def capture(self):
        """Record one iteration (all subsets) into a HIP graph."""
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for k in range(self.nb_subsets):
                self.run_subset(k)
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self, nb_iterations=1):
        if self.graph is None:
            raise RuntimeError("no captured graph for the current schedule: call capture()")
        for _ in range(nb_iterations):
            self.graph.replay()

    def close(self):
        if self.handle:
            self.lib.srfx_tor_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serial_drivers(recons, model, nb_subsets, fused=True, tiled=True):
    """One recon.OSEM per reconstruction (tile plans built here, once)."""
    drivers = [OSEM(model, rc["data"], rc["emap"], nb_subsets=nb_subsets, fused=fused,
                    tiled=tiled) for rc in recons]
    if tiled:
        for d, rc in zip(drivers, recons):
            d.plan(rc["image"])
    return drivers


def serial_iterate(recons, model, nb_subsets, nb_iterations=1, fused=True, tiled=True,
                   drivers=None):
    """The same reconstructions, one after another, through recon.OSEM (the
    plain per-reconstruction path: projection + tiled backprojection)."""
    if drivers is None:
        drivers = serial_drivers(recons, model, nb_subsets, fused, tiled)
    for _ in range(nb_iterations):
        for d, rc in zip(drivers, recons):
            d.subset = 0
            d.iterate(rc["image"], 1)
    return drivers
