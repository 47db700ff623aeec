"""Projection models: the reference's physics layer on the HIP ops.

SplitLoRsModel    TOR model on axis-split LORs (src/python/srf/physics/split_lors.py)
CompleteLoRsModel Siddon model on complete LORs (src/python/srf/physics/complete_lors.py)

The reference transposes the image into each group's frame before every op
and back afterwards, then sums the three groups (split_lors.py:77-113).  Here
each group's op indexes the world image (nx, ny, nz) directly through element
strides, and the three backprojections accumulate into one output image; the
grid, center and size seen by the kernels are the reference's permuted ones,
so the results are the same.
"""
import torch

from . import _lib
from .data import Image, ListModeData, ListModeDataSplit
from .ops import _check_dev, _ptr, _soa, _stream, _vec

# Axis permutation that brings the world image into a group's frame
# (split_lors.py:54-61); the op frame is the permuted image's axes reversed.
PERM = {"x": [0, 2, 1], "y": [1, 2, 0], "z": [2, 1, 0]}


def group_frame(image_shape, center, size, strides, axis):
    """Grid/center/size/strides of the op frame of group ``axis`` for a world
    image of shape (nx, ny, nz) with element ``strides``."""
    p = PERM[axis][::-1]  # op axes x, y, z -> world axes
    grid = [int(image_shape[k]) for k in p]
    cen = [float(center[k]) for k in p]
    siz = [float(size[k]) for k in p]
    st = [int(strides[k]) for k in p]
    return grid, cen, siz, st


# Side streams per device for the concurrent group steps (default priority).
_SIDE_STREAMS = {}


def _concurrent(calls, tensors):
    """Run calls[0] on the current stream and every other call on a side
    stream of its own, each after the work queued so far on the current
    stream; the current stream then waits for all of them (a fork and a join
    through events, so the whole can be captured in a HIP graph).  The
    tensors the calls use are marked as used on the side streams.  Groups of
    one subset run together this way: a group's kernels leave the GPU part
    idle while their last workgroups finish, which the other group's fill."""
    if len(calls) <= 1:
        for c in calls:
            c()
        return
    main = torch.cuda.current_stream()
    side = _SIDE_STREAMS.setdefault(main.device.index, [])
    while len(side) < len(calls) - 1:
        side.append(torch.cuda.Stream(device=main.device))
    fork = torch.cuda.Event()
    fork.record(main)
    joins = []
    for st, c in zip(side, calls[1:]):
        st.wait_event(fork)
        with torch.cuda.stream(st):
            for t in tensors:
                t.record_stream(st)
            c()
        e = torch.cuda.Event()
        e.record(st)
        joins.append(e)
    calls[0]()
    for e in joins:
        main.wait_event(e)


def _world(image):
    data = image.data
    _check_dev(data, "image")
    if data.dim() != 3:
        raise ValueError("image data must be (nx, ny, nz)")
    return data


class SplitLoRsModel:
    """TOR projection model (split_lors.py:28-52).  tof_bin and tof_sigma2
    default to 1e4 as in the reference (TOF stays on: 1e4 < 405758)."""

    AXIS = ("x", "y", "z")

    def __init__(self, kernel_width, geo_sigma2_flag=False, tof_sigma2=None, tof_bin=None):
        self.kernel_width = float(kernel_width)
        self.geo_sigma2_flag = bool(geo_sigma2_flag)
        self.tof_sigma2 = 1.0e4 if tof_sigma2 is None else float(tof_sigma2)
        self.tof_bin = 1.0e4 if tof_bin is None else float(tof_bin)
        # projections read op-frame copies of the image (the reference's
        # transposes) instead of the world image in place
        self.stage_frames = True
        # launch schedule per op: (lanes_per_event, block) or None (default)
        self.schedule = {"projection": None, "backprojection": None, "osem_subset": None}
        # tile shape (slices, wx, wy) and chunk of the tiled plans (make_plans)
        self.tile = None
        self.chunk = None

    def _flags(self, op):
        return _lib.schedule_flags(self.schedule.get(op))

    def autotune(self, image, data, sample=262144, reps=3, ops_=("osem_subset",)):
        """Time every (lanes_per_event, block) schedule of each op in ``ops_``
        on the first ``sample`` events of every group of ``data`` and keep the
        fastest.  Returns {op: {schedule: ms}}."""
        from .autotune import SCHEDULES, _time
        sub = ListModeDataSplit(*[data[a].replace_values(
            None if data[a].values is None else data[a].values[:sample]) for a in self.AXIS])
        for a in self.AXIS:
            sub[a].soa = data[a].soa[:, :sample]
        report = {}
        for op in ops_:
            times = {}
            if op == "backprojection":
                proj = self.projection(image, sub)
            for sch in SCHEDULES:
                self.schedule[op] = sch
                if op == "osem_subset":
                    fn = lambda: self.osem_subset(image, sub)
                elif op == "projection":
                    fn = lambda: self.projection(image, sub)
                else:
                    fn = lambda: self.backprojection(proj, image)
                times[sch] = _time(fn, reps)
            self.schedule[op] = min(times, key=times.get)
            report[op] = times
        return report

    def autotune_tiles(self, image, data, sample=1 << 20, reps=3, tiles=None, chunks=None):
        """Time the tiled OSEM step (projection + tiled backprojection) of the
        first ``sample`` events of the largest group of ``data`` for every
        tile shape x chunk size, on images laid out like ``image``, and keep
        the fastest in self.tile / self.chunk.  Returns {(tile, chunk): ms}."""
        from .autotune import CHUNKS, TILES, _time
        from .plan import TorPlan
        x = _world(image)
        a = max(self.AXIS, key=lambda k: len(data[k]))
        lm = data[a]
        soa = lm.soa[:, :min(sample, len(lm))]
        g, c, sz, st = group_frame(x.shape, image.center, image.size, x.stride(), a)
        out = torch.zeros_like(x)
        times = {}
        for t in tiles or TILES:
            for ch in chunks or CHUNKS:
                plan = TorPlan(soa, g, c, sz, self.kernel_width, self.geo_sigma2_flag,
                               self.tof_bin, self.tof_sigma2, strides=st, tile=t, chunk=ch,
                               axis=a)
                times[(tuple(t), int(ch))] = _time(
                    lambda: plan.osem_subset(x, out=out, accumulate=True,
                                             schedule=self.schedule.get("projection")), reps)
                plan.close()
        best = min(times, key=times.get)
        self.tile, self.chunk = best
        return times

    def _model(self):
        return _lib.make_tor_model(self.kernel_width, self.geo_sigma2_flag, self.tof_bin,
                                   self.tof_sigma2)

    def _grid(self, image, axis):
        data = _world(image)
        g, c, s, st = group_frame(data.shape, image.center, image.size, data.stride(), axis)
        return _lib.make_grid(g, c, s, st)

    def make_plans(self, image, data, ncols=10):
        """One srf_amd.plan.TorPlan per non-empty group of ``data`` for images
        laid out like ``image`` (world (nx, ny, nz), any strides): the events
        binned into image tiles once, for the tiled backprojection /
        Maplors / OSEM step.  Returns {axis: TorPlan or None}."""
        from .plan import DEFAULT_CHUNK, TorPlan, default_tile
        x = _world(image)
        plans = {}
        for a in self.AXIS:
            lm = data[a]
            if len(lm) == 0:
                plans[a] = None
                continue
            g, c, sz, st = group_frame(x.shape, image.center, image.size, x.stride(), a)
            plans[a] = TorPlan(lm.soa, g, c, sz, self.kernel_width, self.geo_sigma2_flag,
                               self.tof_bin, self.tof_sigma2, strides=st, ncols=ncols,
                               tile=self.tile or default_tile(g, sz, self.kernel_width),
                               chunk=self.chunk or DEFAULT_CHUNK, axis=a)
        return plans

    def make_subset_plans(self, image, data, nb_subsets, ncols=10):
        """make_plans for every OSEM subset of ``data`` (contiguous rows of
        each group, recon.subset_bounds), the subsets of every group built
        together (srfx_tor_plan_create_subset_groups).  Returns a list over
        subsets of {axis: TorPlan or None}."""
        from .plan import DEFAULT_CHUNK, TorPlan, default_tile
        from .recon import subset_bounds
        x = _world(image)
        K = int(nb_subsets)
        out = [{a: None for a in self.AXIS} for _ in range(K)]
        groups, axes = [], []
        for a in self.AXIS:
            lm = data[a]
            n = len(lm)
            if n == 0:
                continue
            g, c, sz, st = group_frame(x.shape, image.center, image.size, x.stride(), a)
            groups.append(dict(lors=lm.soa, bounds=[subset_bounds(n, k, K)[0] for k in range(K)]
                               + [n], grid=g, center=c, size=sz, strides=st, axis=a,
                               tile=self.tile or default_tile(g, sz, self.kernel_width)))
            axes.append(a)
        built = TorPlan.subset_groups(groups, self.kernel_width, self.geo_sigma2_flag,
                                      self.tof_bin, self.tof_sigma2, ncols=ncols,
                                      chunk=self.chunk or DEFAULT_CHUNK) if groups else []
        for a, plans in zip(axes, built):
            for k, pl in enumerate(plans):
                if pl.n == 0:
                    pl.close()
                    pl = None
                out[k][a] = pl
        return out

    @staticmethod
    def _check_plan_layout(plan, t):
        if list(plan.strides) != [int(v) for v in group_frame(
                t.shape, plan.center, plan.size, t.stride(), plan.axis)[3]]:
            raise ValueError("image layout differs from the one the plan was built for")

    def _source(self, image, a):
        """The image a projection of group ``a`` reads and its grid: the
        op-frame copy (contiguous, stage_frames) or the world image in place."""
        x = _world(image)
        if self.stage_frames:
            g, c, s, _ = group_frame(x.shape, image.center, image.size, x.stride(), a)
            return self.frame_copy(x, a), _lib.make_grid(g, c, s)
        return x, self._grid(image, a)

    def project_group(self, image, lm, a):
        """Projection of the events ``lm`` (ListModeData) of group ``a``."""
        x = _world(image)
        lors, n, ld = _soa(lm.soa, 10)
        p = torch.empty(n, dtype=torch.float32, device=x.device)
        if n:
            src, g = self._source(image, a)
            _lib.call("srfx_tor_project", _ptr(lors), n, ld, _ptr(src), g, self._model(),
                      _ptr(p), self._flags("projection"), _stream())
        return p

    def projection(self, image, data):
        """split_lors.py:73-90: per group p = Projection(...)."""
        out = {a: data[a].replace_values(self.project_group(image, data[a], a))
               for a in self.AXIS}
        return ListModeDataSplit(out["x"], out["y"], out["z"])

    def backprojection(self, data, image, out=None, plans=None):
        """split_lors.py:93-113: sum over groups of Backprojection(lors_value).
        With ``plans`` (make_plans of the same events) the tiled kernels run."""
        x = _world(image)
        if out is None:
            out = torch.zeros_like(x)
        for a in self.AXIS:
            lm = data[a]
            lors, n, ld = _soa(lm.soa, 10)
            if n == 0:
                continue
            v = _vec(lm.values, n, "lors_value")
            if plans is not None:
                self._planned(plans, a, n, out).backprojection(v, out=out, accumulate=True)
                continue
            _lib.call("srfx_tor_backproject", _ptr(lors), n, ld, _ptr(v),
                      self._grid(Image(out, image.center, image.size), a), self._model(),
                      _ptr(out), _lib.SRFX_ACCUMULATE | self._flags("backprojection"), _stream())
        return Image(out, image.center, image.size)

    def maplors(self, data, image, out=None, plans=None):
        """split_lors.py:116-133: sum over groups of Maplors(lors_value).
        With ``plans`` (make_plans(..., ncols=7) of the same events) the tiled
        kernels run."""
        x = _world(image)
        if out is None:
            out = torch.zeros_like(x)
        for a in self.AXIS:
            lm = data[a]
            lors, n, ld = _soa(lm.soa, 7)
            if n == 0:
                continue
            v = _vec(lm.values, n, "lors_value")
            if plans is not None:
                self._planned(plans, a, n, out).maplors(v, out=out, accumulate=True)
                continue
            _lib.call("srfx_tor_maplors", _ptr(lors), n, ld, _ptr(v),
                      self._grid(Image(out, image.center, image.size), a),
                      self.kernel_width, 1 if self.geo_sigma2_flag else 0, _ptr(out),
                      _lib.SRFX_ACCUMULATE, _stream())
        return Image(out, image.center, image.size)

    @staticmethod
    def _planned(plans, a, n, out):
        plan = plans.get(a)
        if plan is None or plan.n != n:
            raise ValueError(f"no plan for the {n} events of group {a}")
        SplitLoRsModel._check_plan_layout(plan, out)
        return plan

    def osem_subset(self, image, data, out=None, plans=None):
        """Projection + ratio backprojection of all groups (recon_step.py:33-41
        without writing the projections): one fused kernel per group, or with
        ``plans`` (make_plans of the same events) the projection kernel and
        the tiled backprojection of 1/p."""
        x = _world(image)
        if out is None:
            out = torch.zeros_like(x)
        if plans is not None:
            # the groups' steps run concurrently (their scatters add into out
            # with atomics)
            steps = [(a, self._planned(plans, a, len(data[a]), out)) for a in self.AXIS
                     if len(data[a])]
            _concurrent([lambda a=a, pl=pl: self.plan_osem_group(pl, x, out, a)
                         for a, pl in steps], (x, out))
            return Image(out, image.center, image.size)
        for a in self.AXIS:
            lm = data[a]
            lors, n, ld = _soa(lm.soa, 10)
            if n == 0:
                continue
            src, g_in = self._source(image, a)
            g_out = self._grid(Image(out, image.center, image.size), a)
            _lib.call("srfx_tor_osem_subset", _ptr(lors), n, ld, _ptr(src), g_in, g_out,
                      self._model(), _ptr(out), _lib.SRFX_ACCUMULATE | self._flags("osem_subset"),
                      _stream())
        return Image(out, image.center, image.size)

    @staticmethod
    def frame_copy(x, a, out=None):
        """The op frame of group ``a`` of the world image x as a contiguous
        [gz][gy][gx] tensor: the reference's transpose (split_lors.py:79),
        one LDS-tiled pass (srfx_frame_copy)."""
        _check_dev(x, "image")
        g, c, s, st = group_frame(x.shape, (0, 0, 0), (1, 1, 1), x.stride(), a)
        if out is None:
            out = torch.empty((g[2], g[1], g[0]), dtype=torch.float32, device=x.device)
        elif tuple(out.shape) != (g[2], g[1], g[0]) or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous {tuple(g[::-1])} tensor")
        _lib.call("srfx_frame_copy", _ptr(x), _lib.make_grid(g, c, s, st), _ptr(out), _stream())
        return out

    def plan_osem_group(self, plan, x, out, a, stage=None):
        """The OSEM step of one group's plan into the world image ``out``
        (accumulating).  With stage_frames the projection reads the
        op-frame copy of x (a box row of voxels then shares a cache line,
        where the world layout puts each voxel column of the box on its own
        line: 0.57 against 0.79 ms per 1M events at 200x200x100), else x in
        place."""
        self._check_plan_layout(plan, x)
        if stage == "scatter":
            src, op = x, False  # the projection is already in the plan
        elif self.stage_frames:
            src, op = self.frame_copy(x, a), True
        else:
            src, op = x, False
        plan.osem_subset(src, out=out, accumulate=True, schedule=self.schedule.get("projection"),
                         stage=stage, op_frame=op)

    def count_nnz(self, image, data):
        """Total non-zero (event, voxel) pairs over the groups (int64 tensor)."""
        x = _world(image)
        tot = torch.zeros(1, dtype=torch.int64, device=x.device)
        for a in self.AXIS:
            lors, n, ld = _soa(data[a].soa, 10)
            if n:
                _lib.call("srfx_tor_count_nnz", _ptr(lors), n, ld, self._grid(image, a),
                          self._model(), None, _ptr(tot), _stream())
        return tot


class CompleteLoRsModel:
    """Siddon projection model (complete_lors.py:33-56).  The reference
    transposes the image [2,1,0] and passes the transposed grid, center and
    size reversed (complete_lors.py:61-70), so the op's x, y, z are the world
    x, y, z and only the memory order changes; strides express that order."""

    def __init__(self, name="siddon", tof_sigma2=None, tof_bin=None):
        self.name = name
        self.tof_sigma2 = 1.0e4 if tof_sigma2 is None else float(tof_sigma2)
        self.tof_bin = 1.0e4 if tof_bin is None else float(tof_bin)

    @staticmethod
    def _grid(image):
        data = _world(image)
        return _lib.make_grid(list(data.shape), image.center, image.size, list(data.stride()))

    def projection(self, image, data):
        x = _world(image)
        lors, n, ld = _soa(data.soa, 7)
        p = torch.empty(n, dtype=torch.float32, device=x.device)
        if n:
            _lib.call("srfx_siddon_project", _ptr(lors), n, ld, _ptr(x), self._grid(image),
                      self.tof_bin, self.tof_sigma2, _ptr(p), 0, _stream())
        return data.replace_values(p)

    def backprojection(self, data, image, out=None, plans=None):
        """Backprojection (complete_lors.py:72-98); with plans (make_plans of
        the same LORs) through the plan's tiled scatter."""
        x = _world(image)
        if out is None:
            out = torch.zeros_like(x)
        if plans is not None:
            if plans["siddon"] is not None:
                v = _vec(data.values, plans["siddon"].n, "lors_value")
                plans["siddon"].backprojection(v, out=out, accumulate=True)
            return Image(out, image.center, image.size)
        lors, n, ld = _soa(data.soa, 7)
        if n:
            v = _vec(data.values, n, "lors_value")
            _lib.call("srfx_siddon_backproject", _ptr(lors), n, ld, _ptr(v),
                      self._grid(Image(out, image.center, image.size)), self.tof_bin,
                      self.tof_sigma2, _ptr(out), _lib.SRFX_ACCUMULATE, _stream())
        return Image(out, image.center, image.size)

    def osem_subset(self, image, data, out=None, plans=None):
        """Projection + ratio backprojection (recon_step.py:33-41 over
        complete_lors.py:59-98) in one fused kernel (srfx_siddon_osem_subset):
        out += A^T (1 / A x), events with A x = 0 left out as the reference's
        backprojection does (siddon.cu.cc:363-371).  With plans (make_plans
        of the same LORs), the projection kernel, then the tiled
        backprojection of the plan."""
        x = _world(image)
        if out is None:
            out = torch.zeros_like(x)
        if plans is not None:
            if plans["siddon"] is not None:
                plans["siddon"].osem_subset(x, out=out, accumulate=True)
            return Image(out, image.center, image.size)
        lors, n, ld = _soa(data.soa, 7)
        if n:
            _lib.call("srfx_siddon_osem_subset", _ptr(lors), n, ld, _ptr(x), self._grid(image),
                      self.tof_bin, self.tof_sigma2, _ptr(out), _lib.SRFX_ACCUMULATE, _stream())
        return Image(out, image.center, image.size)

    def make_plans(self, image, data, tile=None, chunk=None):
        """Tiled Siddon plan of data's LORs for images laid out like image
        (SiddonPlan), as {"siddon": plan} (None without LORs)."""
        from .plan import DEFAULT_SIDDON_CHUNK, DEFAULT_SIDDON_TILE, SiddonPlan
        x = _world(image)
        if data.soa.shape[1] == 0:
            return {"siddon": None}
        return {"siddon": SiddonPlan(data.soa, list(x.shape), image.center, image.size,
                                     self.tof_bin, self.tof_sigma2, strides=x.stride(),
                                     tile=tile or DEFAULT_SIDDON_TILE,
                                     chunk=chunk or DEFAULT_SIDDON_CHUNK)}


    def make_subset_plans(self, image, data, nb_subsets, tile=None, chunk=None):
        """make_plans for every OSEM subset of ``data`` (contiguous rows,
        recon.subset_bounds), built together (srfx_siddon_plan_create_subsets).
        Returns a list over subsets of {"siddon": SiddonPlan}."""
        from .plan import DEFAULT_SIDDON_CHUNK, DEFAULT_SIDDON_TILE, SiddonPlan
        from .recon import subset_bounds
        x = _world(image)
        n = len(data)
        K = int(nb_subsets)
        bounds = [subset_bounds(n, k, K)[0] for k in range(K)] + [n]
        plans = SiddonPlan.subsets(data.soa, bounds, list(x.shape), image.center, image.size,
                                   self.tof_bin, self.tof_sigma2, strides=x.stride(),
                                   tile=tile or DEFAULT_SIDDON_TILE,
                                   chunk=chunk or DEFAULT_SIDDON_CHUNK)
        return [{"siddon": pl} for pl in plans]

    def maplors(self, data, image, out=None):
        x = _world(image)
        if out is None:
            out = torch.zeros_like(x)
        lors, n, ld = _soa(data.soa, 6)
        if n:
            v = _vec(data.values, n, "lors_value")
            _lib.call("srfx_siddon_maplors", _ptr(lors), n, ld, _ptr(v),
                      self._grid(Image(out, image.center, image.size)), _ptr(out),
                      _lib.SRFX_ACCUMULATE, _stream())
        return Image(out, image.center, image.size)


def projection(model, image, data):
    """model/_projection.py:37-40 dispatch."""
    return model.projection(image, data)


def backprojection(model, data, image):
    """model/_backprojection.py:37-41 dispatch."""
    return model.backprojection(data, image)
