"""TOF OSEM throughput benchmark on one MI355X.

Workload (tof_tor_c2): TOF tube-of-response list-mode OSEM, 200x200x100 image
of 2.0x2.0x2.2 mm, 20M synthetic events (cylinder + 6 spheres, 530 ps FWHM,
kernel_width 4 mm), 10 subsets.  A step is one OSEM iteration: every event is
projected and backprojected once.  Per subset and axis group: the projection
kernel, then the tiled backprojection of 1/p (events binned into image tiles
once per subset, accumulated in LDS), the axis groups' kernels on their own
streams at once; then the image update.  Events are
generated, preprocessed (axis groups, locality order inside each subset), the
tile plans built and the efficiency map computed on the GPU before the timed
region; inputs are resident in HBM when it starts.  --untiled times the
single fused projection + scatter kernel per group instead.

Prints one JSON line:  value = events * steps / time (Mevents/s, proj+bp), the
time being the timed iterations plus steps/3 of the tile-plan build (the
workload's reconstruction runs 3 iterations over one event set, and the plans
are built once per set; value_excl_plan_build leaves the build out);
roofline = the dominant kernel's algorithmic bytes per launch over its mean
launch time (HIP events on the launch stream, in as many iterations again with
each group's kernels launched alone); cpu_baseline = the FP32 CPU
port of the reference kernels on a bounded sample of the same events.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
C2_ITERATIONS = 3  # iterations of the TOF OSEM workload's reconstruction
LOR_BYTES = 40  # 10 float32 columns per TOR LOR record


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=int, default=20_000_000)
    ap.add_argument("--subsets", type=int, default=10)
    ap.add_argument("--effmap-pairs", type=int, default=0,
                    help="crystal pairs for the efficiency map (0 = all pairs)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target wall time of the CPU baseline sample (0 = skip)")
    ap.add_argument("--unfused", action="store_true",
                    help="time separate projection + backprojection kernels")
    ap.add_argument("--workload", default="tof_tor_c2",
                    choices=["tof_tor_c2", "tof_tor_c2_kw8", "siddon_c1", "autograd_c3",
                             "batched_c4", "large_c5", "hbm_copy", "psf_c2", "io_c2"],
                    help="tof_tor_c2_kw8: the C2 workload at kernel_width 8 mm (every "
                         "event's candidate box exceeds 4 voxels: the chord-walk kernels)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the projection schedule and the tile shape on a sample of the "
                         "events before timing (default: the fixed defaults, tile 12x32x10, "
                         "chunk 2048; the tuner's candidates lie within a few percent of "
                         "each other, inside its timing noise)")
    ap.add_argument("--no-sort", dest="sort", action="store_false",
                    help="keep the generated event order inside each subset (default: "
                         "reorder events inside each subset for locality)")
    ap.add_argument("--untiled", action="store_true",
                    help="one fused projection + scatter kernel per group instead of the "
                         "projection kernel + tiled backprojection")
    return ap.parse_args()


def _measured_traffic(workload):
    """HBM bytes per launch of each kernel of ``workload`` from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE scaled by a plain
    copy; see profiles/), keyed by op name, or {} if there is none."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"hbm_traffic_{workload}.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return {}
    out = {}
    for name, k in d.get("kernels", {}).items():
        out[name] = {"bytes_per_launch": int(k["bytes_per_launch"]),
                     "source": f"profiles/hbm_traffic_{workload}.json ({k['kernel']})"}
    return out


def _path_label(args, tiled):
    """The kernels one timed subset step runs, per axis group."""
    if args.unfused:
        return ("per group: tor_kernel<kProj> projection, then tor_kernel<kBack> ratio "
                "backprojection (FP32 global atomics); mlem_update per subset")
    if tiled:
        return ("per group, the groups of a subset on their own streams at once: "
                "frame_copy_kernel (op-frame copy of the image), "
                "record_proj_kernel (projection into the plan's event records), "
                "tile_scatter<kBack> (tiled backprojection of 1/p); mlem_update per subset")
    return ("per group: frame_copy_kernel, tor_kernel<kOsem> (fused projection + ratio "
            "backprojection, FP32 global atomics); mlem_update per subset")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def lscpu_topology():
    """Sockets, cores per socket and threads per core of the host, from lscpu
    (None where it does not say)."""
    import subprocess
    out = {}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    keys = {"Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "Model name": "model"}
    for line in text.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys:
            v = v.strip()
            out[keys[k.strip()]] = int(v) if v.isdigit() else v
    return out or None


def cpu_baseline(lors_dev, meta, target_s):
    """FP32 CPU port of tor.cu.cc (projection + ratio backprojection) on the
    first events of the x group (a device tensor, copied to the host once),
    as the CPU baseline plan says: a short sizing run, then on a sample sized
    so that one projection + backprojection takes about target_s / 4, one
    warm-up and the median of three timed runs of each op."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import lib as O
    O.build()
    threads = len(os.sched_getaffinity(0))
    g = meta["frame_x"]
    args = (g["grid"], g["center"], g["size"], meta["kernel_width"], False, meta["tof_bin"],
            meta["tof_sigma2"])
    ones = np.ones(int(np.prod(g["grid"])), np.float32)
    total = lors_dev.shape[1]

    def run(sub):
        t0 = time.perf_counter()
        p = O.tor_project(sub, ones, *args, prec="f32", nthreads=threads)
        t1 = time.perf_counter()
        O.tor_backproject(sub, p.astype(np.float32), g["grid"], g["center"], g["size"],
                          meta["kernel_width"], False, meta["tof_bin"], meta["tof_sigma2"],
                          prec="f32", nthreads=threads)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    n = min(20000, total)
    tp, tb = run(np.ascontiguousarray(lors_dev[:, :n].cpu().numpy()))
    n = int(min(total, max(n, n * (target_s / 4) / max(tp + tb, 1e-3))))
    sub = np.ascontiguousarray(lors_dev[:, :n].cpu().numpy())
    run(sub)  # warm-up
    runs = [run(sub) for _ in range(3)]
    tp = float(np.median([r[0] for r in runs]))
    tb = float(np.median([r[1] for r in runs]))
    return {"value": n / (tp + tb) / 1e6, "unit": "Mevents/s", "cores": threads,
            "kind": "port", "lscpu": lscpu_topology(),
            "projection_s": round(tp, 3), "backprojection_s": round(tb, 3),
            "sample": f"first {n} events of the x axis group; 1 warm-up, median of 3 runs: "
                      f"projection {tp:.2f}s + ratio backprojection {tb:.2f}s; FP32 port of "
                      f"tor.cu.cc (all {g['grid'][2]} slices tested per event, as the "
                      f"reference does), OpenMP over {threads} threads (the process affinity "
                      f"mask)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU runs are not "
                          "part of this build (use --gpus 1)"}))
        sys.exit(2)
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    if args.workload not in ("tof_tor_c2", "tof_tor_c2_kw8"):
        import bench_workloads
        bench_workloads.run(args)
        return
    import torch
    from srf_amd import _lib, effmap, physics, synth, preprocess
    from srf_amd.data import Image, split_from_groups
    from srf_amd.plan import DEFAULT_CHUNK
    from srf_amd.recon import OSEM, subset_view
    from srf_amd.scanner import CylindricalScanner

    dev = torch.device("cuda:0")
    torch.cuda.init()
    scanner = CylindricalScanner()
    t0 = time.time()
    ph, ev, meta = synth.c2(n_events=args.events, device=dev)
    if args.workload == "tof_tor_c2_kw8":
        meta["kernel_width"] = 8.0
    torch.cuda.synchronize()
    log(f"generated {args.events} events in {time.time() - t0:.1f}s")
    t0 = time.time()
    groups = preprocess.recon_process(ev, scanner.tof_res_ps)
    del ev
    torch.cuda.synchronize()
    log(f"preprocessed in {time.time() - t0:.1f}s: " +
        ", ".join(f"{a}={groups[a].shape[1]}" for a in "xyz"))
    grid, center, size = meta["grid"], meta["center"], meta["size"]
    model = physics.SplitLoRsModel(meta["kernel_width"], meta["geo_sigma2_flag"],
                                   tof_sigma2=meta["tof_sigma2"], tof_bin=meta["tof_bin"])
    t0 = time.time()
    npairs = args.effmap_pairs or None
    summed = effmap.compute_tor_effmap(scanner, model, grid, center, size, dev,
                                       max_pairs=npairs)
    E = effmap.finalize(summed, grid, center, size, scanner.inner_radius, scanner.axial_length)
    torch.cuda.synchronize()
    log(f"efficiency map over {npairs or scanner.nb_pairs()} crystal pairs in "
        f"{time.time() - t0:.1f}s")
    x = effmap.initial_image(grid, center, size, scanner.inner_radius, scanner.axial_length, dev)
    image = Image(x, center, size)
    emap = Image(E, center, size)
    if args.sort:
        groups, _ = preprocess.reorder_groups(groups, args.subsets)
    data = split_from_groups(groups)
    tuned, tuned_tiles = None, None
    if args.autotune:
        t0 = time.time()
        if args.untiled:
            tuned = model.autotune(image, data, ops_=("osem_subset",))
        else:
            tuned = model.autotune(image, data, ops_=("projection",))
            tuned_tiles = model.autotune_tiles(image, data)
        log(f"autotuned schedules {model.schedule}, tile {model.tile}, chunk {model.chunk} in "
            f"{time.time() - t0:.1f}s")
    elif not args.untiled:
        # the tiled step's projection: 4 lanes per event, 256 threads (0.77 ms per
        # 1M events at this workload against 0.92 ms at 1 lane per event)
        model.schedule["projection"] = (4, 256)
    osem = OSEM(model, data, emap, nb_subsets=args.subsets, fused=not args.unfused,
                tiled=not args.untiled)
    plan_s, plan_cold_s, listings, plan_runs_ms = None, None, None, None
    if osem.tiled:
        # the first build in a process also pays one-time costs (first
        # launches, first mapping of the plans' memory), as the first
        # iteration does: like the iterations, the timed builds follow a
        # warm-up build, whose figure is reported beside them; the build
        # charged is the median of three (a build's allocations and host
        # round trips vary by ~1 ms from run to run)
        builds = []
        for _ in range(4):
            osem.close()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            osem.plan(image)
            torch.cuda.synchronize()
            builds.append(time.perf_counter() - t0)
        plan_cold_s, plan_s = builds[0], sorted(builds[1:])[1]
        plan_runs_ms = [round(1e3 * b, 2) for b in builds[1:]]
        listings = sum(p.info()[0] for ps in osem.plans for p in ps.values() if p is not None)
        log(f"tile plans for {args.subsets} subsets in {plan_s:.2f}s: "
            f"{listings / args.events:.1f} (event, tile) listings per event")

    # exact non-zero counts per launch (x, y, z group of each subset)
    nnz = {}
    for k, sub in enumerate(osem.subsets):
        for a in "xyz":
            one = split_from_groups({b: (sub[b].soa if b == a else sub[b].soa[:, :0])
                                     for b in "xyz"})
            nnz[(k, a)] = int(model.count_nnz(image, one).item())
    n_launch = {(k, a): len(osem.subsets[k][a]) for k in range(args.subsets) for a in "xyz"}
    total_nnz = sum(nnz.values())
    log(f"nnz per event {total_nnz / args.events:.1f}")

    # per-launch HIP events around each group's kernels (torch's current
    # stream is the launch stream)
    timings = {}
    orig = model.osem_subset

    def timed_tiled(img, sub, out=None, plans=None):
        # the two kernels of srfx_tor_plan_osem_subset, launched one by one
        # (projection into the plan's event records, then the tiled scatter)
        for a in "xyz":
            if len(sub[a]) == 0:
                continue
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e0.record()
            # the projection half includes the copy of the image into the
            # group's op frame
            model.plan_osem_group(plans[a], img.data, out, a, stage="project")
            e1.record()
            model.plan_osem_group(plans[a], img.data, out, a, stage="scatter")
            e2.record()
            timings.setdefault(a, []).append((e0, e1, e2))
        return Image(out, img.center, img.size)

    def timed_osem(img, sub, out=None):
        for a in "xyz":
            if len(sub[a]) == 0:
                continue
            only = physics.ListModeDataSplit(*[sub[b] if b == a else sub[b].replace_values(None)
                                               for b in "xyz"])
            for b in "xyz":
                if b != a:
                    only[b].soa = sub[b].soa[:, :0]
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            orig(img, only, out=out)
            e1.record()
            timings.setdefault(a, []).append((e0, e1, sub[a].soa.data_ptr()))
        return Image(out, img.center, img.size)

    for _ in range(args.warmup):
        osem.iterate(image, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s0.record()
    osem.iterate(image, args.steps)
    s1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = s0.elapsed_time(s1)
    # per-launch kernel times: as many iterations again, each group's kernels
    # launched alone between HIP events (the timed iterations above run the
    # groups of a subset at once)
    if osem.tiled and not args.unfused:
        model.osem_subset = timed_tiled
    elif not args.unfused:
        model.osem_subset = timed_osem
    timings.clear()
    osem.iterate(image, args.steps)
    torch.cuda.synchronize()
    model.osem_subset = orig
    # The tile plans are built once per event set, and this workload's
    # reconstruction runs C2_ITERATIONS iterations over one set: each timed
    # iteration carries 1/C2_ITERATIONS of the build (the reference builds
    # nothing, so the comparison must include it).
    plan_share_ms = 1e3 * plan_s * args.steps / C2_ITERATIONS if plan_s is not None else 0.0
    ms_per_step = (gpu_ms + plan_share_ms) / args.steps
    value = args.events * args.steps / ((gpu_ms + plan_share_ms) / 1e3) / 1e6
    value_kernels = args.events * args.steps / (gpu_ms / 1e3) / 1e6
    log(f"{args.steps} iterations: {gpu_ms:.1f} ms (wall {wall * 1e3:.1f} ms), "
        f"{value_kernels:.1f} Mevents/s; with the plan build amortised over "
        f"{C2_ITERATIONS} iterations {value:.1f} Mevents/s")
    assert torch.isfinite(image.data).all().item()

    roof = None
    kernels = None
    nvox = int(np.prod(grid))
    if timings and osem.tiled:
        # dominant kernel: the slower of the x group's projection and tiled
        # backprojection (the z group is empty here)
        a = max(timings, key=lambda k: sum(e0.elapsed_time(e2) for e0, _, e2 in timings[k]))
        nb = len(timings[a])
        launches = [(n_launch[(i % args.subsets, a)], nnz[(i % args.subsets, a)])
                    for i in range(nb)]
        byts = float(np.mean([n * (LOR_BYTES + 4) + 4 * z + 4 * nvox for n, z in launches]))
        kernels = {}
        fr = physics.group_frame(grid, center, size, (1, 1, 1), a)
        box = _lib.tor_patch_info(fr[0], fr[2], meta["kernel_width"])[1]
        walk = (f"candidate-box walk, box {box}" if box else "chord walk (box wider than 4)")
        # every event of the group's plans on the box-by-(box - 1) walk: the
        # scatter runs its single-walk kernel
        single = box and all(ps[a].walks()[1] for ps in osem.plans if ps.get(a) is not None)
        swalk = f"{walk}, single-walk kernel" if single else walk
        for name, sel, kname in (("projection", lambda t: t[0].elapsed_time(t[1]),
                                  f"record_proj_kernel ({walk} from the walk records, values "
                                  "into their value slots)"),
                                 ("backprojection", lambda t: t[1].elapsed_time(t[2]),
                                  f"tile_scatter<kBack> ({swalk}; FP64 LDS tile accumulation, "
                                  "coalesced flush)")):
            ms = float(np.mean([sel(t) for t in timings[a]]))
            ach = byts / (ms / 1e3) / 1e9
            kernels[name] = {"kernel": kname, "group": a, "mean_launch_ms": round(ms, 4),
                             "algorithmic_bytes_per_launch": int(byts),
                             "achieved_GBps": round(ach, 1),
                             "frac": round(ach / HBM_PEAK_GBS, 4)}
        traffic = _measured_traffic(args.workload)
        for name in kernels:
            t = traffic.get(name)
            if t:
                kernels[name]["traffic_bytes_per_launch"] = t["bytes_per_launch"]
        dom = max(kernels, key=lambda k: kernels[k]["mean_launch_ms"])
        k = kernels[dom]
        tr = traffic.get(dom)
        roof = {"bound": "hbm", "achieved": k["achieved_GBps"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": k["frac"],
                "traffic": tr["bytes_per_launch"] if tr else None,
                "traffic_source": tr["source"] if tr else None, "kernel": k["kernel"],
                "group": a, "mean_launch_ms": k["mean_launch_ms"],
                "algorithmic_bytes_per_launch": k["algorithmic_bytes_per_launch"],
                "bytes_formula": "sum_i (40 + 4 + 4*nnz_i) + 4*N_vox per op"}
    elif timings:
        # dominant kernel: the x group's fused pass (the z group is empty here)
        a = max(timings, key=lambda k: sum(e0.elapsed_time(e1) for e0, e1, _ in timings[k]))
        durs = [e0.elapsed_time(e1) for e0, e1, _ in timings[a]]
        per_sub = [(n_launch[(k, a)], nnz[(k, a)]) for k in range(args.subsets)]
        nb = len(durs)
        launches = [per_sub[i % args.subsets] for i in range(nb)]
        byts = [n * (LOR_BYTES + 4) + 8 * z + 2 * 4 * nvox for n, z in launches]
        mean_bytes = float(np.mean(byts))
        mean_ms = float(np.mean(durs))
        achieved = mean_bytes / (mean_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "tor_kernel<kOsem> (fused projection + ratio backprojection)",
                "group": a, "mean_launch_ms": round(mean_ms, 4),
                "algorithmic_bytes_per_launch": int(mean_bytes),
                "bytes_formula": "sum_i (40 + 4 + 8*nnz_i) + 2*4*N_vox"}

    cpu = None
    if args.cpu_seconds > 0:
        f = physics.group_frame(grid, center, size, (1, 1, 1), "x")
        meta["frame_x"] = {"grid": f[0], "center": f[1], "size": f[2]}
        try:
            cpu = cpu_baseline(groups["x"], meta, args.cpu_seconds)
        except Exception as exc:  # the baseline must not hide the GPU result
            cpu = {"error": repr(exc)}

    out = {
        "metric": "TOF OSEM Mevents/s (proj+backproj) + % HBM roofline; batched/grad vs plain",
        "value": round(value, 2), "unit": "Mevents/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "value_excl_plan_build": round(value_kernels, 2),
        "plan_build_ms": round(plan_s * 1e3, 2) if plan_s is not None else None,
        "plan_build_cold_ms": round(plan_cold_s * 1e3, 2) if plan_cold_s is not None else None,
        "plan_build_runs_ms": plan_runs_ms,
        "plan_build_amortised_over_iterations": C2_ITERATIONS if plan_s is not None else None,
        "config": {"workload": args.workload, "image": list(grid), "voxel_mm": list(meta["voxel"]),
                   "events": args.events, "subsets": args.subsets,
                   "step": "one OSEM iteration over all events",
                   "kernel_width_mm": meta["kernel_width"], "tof_fwhm_ps": 530,
                   "tof_sigma2_mm2": round(meta["tof_sigma2"], 3), "tof_bin_mm": meta["tof_bin"],
                   "nnz_per_event": round(total_nnz / args.events, 2),
                   "path": _path_label(args, osem.tiled), "sorted": bool(args.sort),
                   "tiled": osem.tiled,
                   "tile": (list(next(p.tile for ps in osem.plans for p in ps.values()
                                      if p is not None)) if osem.tiled else None),
                   "chunk_default": None if model.chunk else DEFAULT_CHUNK,
                   "plan_build_s": round(plan_s, 3) if plan_s is not None else None,
                   "listings_per_event": (round(listings / args.events, 2)
                                          if listings is not None else None),
                   "schedule": model.schedule["osem_subset" if args.untiled else "projection"],
                   "chunk": model.chunk,
                   "autotune_ms": ({op: {str(k): round(v, 3) for k, v in t.items()}
                                    for op, t in tuned.items()} if tuned else None),
                   "autotune_tiles_ms": ({str(k): round(v, 3) for k, v in tuned_tiles.items()}
                                         if tuned_tiles else None)},
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
