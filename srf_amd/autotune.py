"""Launch-schedule autotuner for the TOR kernels.

Tile plans are tuned separately (SplitLoRsModel.autotune_tiles): tile shape
(slices x wx x wy, TILES) and events per workgroup (CHUNKS), timed on the
OSEM step of a sample of the caller's own events.

A schedule is (lanes_per_event, block): how many lanes share one event (its
crossed slices interleaved between them: 1, 4 or 16) and the LOR tile, i.e.
threads per block (64, 128 or 256).  ``tune`` times every schedule of an op on
a bounded sample of the caller's own events with HIP events on the current
stream and keeps the fastest; results are cached per (op, patch size, frame
shape, kernel width, TOF on/off) for the life of the process and can be saved
to / loaded from JSON.  Results of the ops agree to FP32 rounding across
schedules, so tuning never changes what is computed.
"""
import json
import math

import torch

from . import ops

LANES = (1, 4, 16)
BLOCKS = (64, 128, 256)
SCHEDULES = tuple((l, b) for l in LANES for b in BLOCKS)
# Tile plans: (slices, wx, wy) and events per workgroup chunk.
TILES = ((12, 32, 10), (12, 32, 12), (8, 32, 20), (16, 32, 10), (8, 32, 16), (8, 16, 32),
         (8, 20, 32), (10, 32, 16), (4, 32, 32), (16, 16, 16), (8, 32, 32))
CHUNKS = (1024, 2048, 4096)

_cache = {}


def patch_size(grid, size, kernel_width):
    """TOR patch P = ceil((sqrt(2) kw + pitch_z) / pitch_x) (tor.cu.cc:375)."""
    ix = size[0] / grid[0]
    iz = size[2] / grid[2]
    return int(math.ceil((math.sqrt(2.0) * kernel_width + iz) / ix))


def key(op, grid, size, kernel_width, tof):
    return (op, patch_size(grid, size, kernel_width), tuple(int(g) for g in grid),
            round(float(kernel_width), 6), bool(tof))


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def tune(op, lors, image, grid, center, size, kernel_width, geo_sigma2_flag, tof_bin,
         tof_sigma2, sample=262144, reps=3, schedules=SCHEDULES):
    """Fastest schedule for ``op`` in {'projection', 'backprojection',
    'osem_subset'} on the first ``sample`` events of ``lors`` ([10, N] in the
    op frame; image in the op frame).  Returns (schedule, {schedule: ms})."""
    k = key(op, grid, size, kernel_width, tof_sigma2 < 405758.0)
    if k in _cache:
        return _cache[k], None
    L = lors[:, :min(sample, lors.shape[1])]
    args = (grid, center, size)
    if op == "backprojection":
        v = ops.projection(L, image, *args, kernel_width, geo_sigma2_flag, tof_bin, tof_sigma2)
    times = {}
    for sch in schedules:
        if op == "projection":
            fn = lambda: ops.projection(L, image, *args, kernel_width, geo_sigma2_flag, tof_bin,
                                        tof_sigma2, schedule=sch)
        elif op == "backprojection":
            fn = lambda: ops.backprojection(image, L, v, *args, tof_bin, tof_sigma2, kernel_width,
                                            geo_sigma2_flag, schedule=sch)
        elif op == "osem_subset":
            fn = lambda: ops.osem_subset(L, image, *args, kernel_width, geo_sigma2_flag, tof_bin,
                                         tof_sigma2, schedule=sch)
        else:
            raise ValueError(op)
        times[sch] = _time(fn, reps)
    best = min(times, key=times.get)
    _cache[k] = best
    return best, times


def lookup(op, grid, size, kernel_width, tof):
    return _cache.get(key(op, grid, size, kernel_width, tof))


def save(path):
    with open(path, "w") as f:
        json.dump([[list(k), list(v)] for k, v in _cache.items()], f)


def load(path):
    with open(path) as f:
        for k, v in json.load(f):
            k[2] = tuple(k[2])
            _cache[tuple(k)] = tuple(v)
