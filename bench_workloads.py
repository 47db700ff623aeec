"""Secondary benchmark workloads of bench.py (``--workload``); each prints one
JSON line in bench.py's format.

siddon_c1    non-TOF Siddon projection + backprojection, 2 iterations x 4
             subsets of list-mode OSEM with Poisson counts; 128x128x64 image
             (3.0x3.0x3.4 mm), 1M random crystal pairs.  A step is one OSEM
             iteration.
autograd_c3  differentiable projector: batch of 8 lumpy 128x128x64 images,
             2M TOF events each, Poisson NLL; forward + backward through
             torch.autograd against the plain projection + adjoint calls, and
             the memory of recomputation against stored voxel lists.
batched_c4   64 independent TOF OSEM reconstructions (image sizes mixed from
             96x96x48, 128x128x64, 200x200x100; events log-uniform in
             [1e5, 5e6]), 2 iterations x 8 subsets: one persistent launch per
             subset replayed from a HIP graph, against the same
             reconstructions run one after another.
large_c5     256x256x160 (1.6x1.6x1.7 mm) torso phantom, 300M events at
             400 ps, attenuated efficiency map, 2 iterations x 20 subsets.
psf_c2       the PSF blur pair of one PSF-modelled OSEM step (forward blur
             of the image, backward blur of the backprojection) on the
             200x200x100 image, Gaussian kernels (4 mm transaxial, 4.4 mm
             axial FWHM).
io_c2        20M list-mode rows (float32 N x 7 .npy, 560 MB) streamed from a
             file to device memory through the pinned double buffer.
"""
import json
import os
import sys
import time

import numpy as np


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def _base(metric, value, unit, steps, warmup, ms_per_step, config, **extra):
    out = {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic", "config": config}
    out.update(extra)
    return out


def _timer():
    import torch
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    return a, b


def run(args):
    {"siddon_c1": siddon_c1, "autograd_c3": autograd_c3, "batched_c4": batched_c4,
     "large_c5": large_c5, "hbm_copy": hbm_copy, "psf_c2": psf_c2,
     "io_c2": io_c2}[args.workload](args)


# ------------------------------------------------------------ calibration ---

def hbm_copy(args):
    """A plain 1 GiB device-to-device copy (torch's copy kernel): the reference
    point that scales the FETCH_SIZE / WRITE_SIZE counters of a profiled
    kernel to bytes (the counters of this copy should read 1 GiB each)."""
    import torch
    dev = torch.device("cuda:0")
    n = 1 << 28  # 2^28 float32 = 1 GiB
    src = torch.rand(n, device=dev)
    dst = torch.empty_like(src)
    for _ in range(max(args.warmup, 1)):
        dst.copy_(src)
    torch.cuda.synchronize()
    a, b = _timer()
    a.record()
    for _ in range(args.steps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    gbs = 2 * 4 * n / (ms / 1e3) / 1e9
    print(json.dumps(_base("device copy GB/s (read + write)", round(gbs, 1), "GB/s", args.steps,
                           args.warmup, round(ms, 4),
                           {"workload": "hbm_copy", "bytes_read": 4 * n, "bytes_written": 4 * n},
                           dtype="f32")), flush=True)


# --------------------------------------------------------------------- C1 ---

# C1 and C5 run 2 OSEM iterations: the plan build is charged at half per
# iteration
C1_ITERATIONS = 2
C5_ITERATIONS = 2


def siddon_c1(args):
    import torch
    from srf_amd import effmap, physics, synth
    from srf_amd.data import Image, ListModeData
    from srf_amd.scanner import CylindricalScanner
    dev = torch.device("cuda:0")
    sc = CylindricalScanner()
    grid, center, size = (128, 128, 64), (0.0, 0.0, 0.0), (384.0, 384.0, 217.6)
    ph = synth.c1_phantom(seed=1)
    truth = torch.from_numpy(ph.voxelize(grid, center, size)).to(dev)
    pairs = torch.from_numpy(synth.c1_pairs(1_000_000, seed=1, scanner=sc)).to(dev)
    soa = torch.zeros(7, pairs.shape[0], device=dev)
    soa[0:6] = pairs.t()
    model = physics.CompleteLoRsModel("siddon", tof_sigma2=1.0e6, tof_bin=1.0)
    p_true = model.projection(Image(truth, center, size), ListModeData(soa, soa=True)).values
    alpha = 10.0 / float(p_true.mean())
    gen = torch.Generator(device=dev)
    gen.manual_seed(1)
    counts = torch.poisson(alpha * p_true, generator=gen)
    keep = counts > 0
    soa, counts = soa[:, keep].contiguous(), counts[keep]
    n = soa.shape[1]
    log(f"C1: {n} LORs with counts (alpha {alpha:.3g})")
    # sensitivity: Siddon Maplors over all crystal pairs, then merge_map's finalisation
    sens = torch.zeros(grid, device=dev)
    total = sc.nb_pairs()
    for s0 in range(0, total, 8_000_000):
        pr = torch.from_numpy(sc.all_pairs((s0, min(total, s0 + 8_000_000)))).to(dev)
        ps = torch.zeros(7, pr.shape[0], device=dev)
        ps[0:6] = pr.t()
        model.maplors(ListModeData(ps, torch.ones(pr.shape[0], device=dev), soa=True),
                      Image(sens, center, size), out=sens)
    E = effmap.finalize(sens, grid, center, size, sc.inner_radius, sc.axial_length)
    x = effmap.initial_image(grid, center, size, sc.inner_radius, sc.axial_length, dev)
    K = 4
    bounds = [((k * n) // K, ((k + 1) * n) // K) for k in range(K)]
    bp = torch.empty_like(x)
    from srf_amd import ops
    from srf_amd.plan import SiddonPlan
    st = tuple(x.stride())
    # tiled backprojection: one plan per subset, built once per event set
    # and charged to the headline spread over the C1 iterations.  The first
    # build in a process also pays one-time costs (first launches of the
    # build's kernels, first mapping of its memory), as the first iteration
    # does; like the iterations, the timed build is the one after a warm-up
    # build, and the cold figure is reported beside it.
    plan_times = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plans = SiddonPlan.subsets(soa, [lo for lo, _ in bounds] + [n], grid, center, size,
                                   1.0, 1.0e6, strides=st)
        torch.cuda.synchronize()
        plan_times.append(time.perf_counter() - t0)
        if len(plan_times) < 2:
            for pl in plans:
                pl.close()
    plan_s = plan_times[1]
    subs = [ListModeData(soa[:, lo:hi], soa=True) for lo, hi in bounds]

    cnts = [counts[lo:hi].contiguous() for lo, hi in bounds]

    def iteration():
        # each subset: the plan's projection over its sorted LOR copy, then
        # the tiled backprojection of c / p (lors_value = p / c,
        # sum_i c_i a_ij / p_i), then the update
        for pl, c in zip(plans, cnts):
            pl.osem_subset(x, out=bp, counts=c)
            ops.mlem_update(x, E, bp)

    for _ in range(args.warmup):
        iteration()
    torch.cuda.synchronize()
    a, b = _timer()
    a.record()
    for _ in range(args.steps):
        iteration()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    value_excl = n / (ms / 1e3) / 1e6
    value = n / ((ms + plan_s * 1e3 / C1_ITERATIONS) / 1e3) / 1e6
    # per-kernel timing for the roofline on the first subset: the projection
    # kernel, the tiled backprojection, and the untiled one for comparison
    lo, hi = bounds[0]
    lm = subs[0]
    # the Siddon op frame is the world frame (complete_lors.py:61-70)
    per, tot = ops.siddon_count_nnz(lm.soa, grid, center, size, 1.0, 1.0e6)
    nnz = int(tot.item())
    v0 = model.projection(Image(x, center, size), lm).values / counts[lo:hi]

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        a, b = _timer()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    pms = timed(lambda: plans[0].osem_subset(x, stage="project"))
    bms = timed(lambda: plans[0].osem_subset(None, out=bp, counts=cnts[0], stage="scatter"))
    ums = timed(lambda: model.backprojection(lm.replace_values(v0), Image(x, center, size),
                                             out=bp.zero_()))
    nl = hi - lo
    nruns, nchunks = plans[0].info()
    byts = nl * (28 + 4) + 4 * nnz + 4 * x.numel()
    kern = {}
    for name, kname, kms in (("projection", "siddon_kernel<kProj> (one lane per LOR, sorted "
                              "LOR stream)", pms),
                             ("backprojection", "sid_tile_scatter (per-tile runs, FP64 LDS "
                              "tile, coalesced flush)", bms)):
        ach = byts / (kms / 1e3) / 1e9
        kern[name] = {"kernel": kname, "mean_launch_ms": round(kms, 4),
                      "algorithmic_bytes_per_launch": byts, "achieved_GBps": round(ach, 1),
                      "frac": round(ach / 8000.0, 4)}
    from bench import _measured_traffic
    traffic = _measured_traffic("siddon_c1")
    for name in kern:
        if traffic.get(name):
            kern[name]["traffic_bytes_per_launch"] = traffic[name]["bytes_per_launch"]
    dom = max(kern, key=lambda k: kern[k]["mean_launch_ms"])
    for pl in plans:
        pl.close()
    cpu = None
    if args.cpu_seconds > 0:
        from bench import lscpu_topology
        from oracle import lib as O
        O.build()
        thr = len(os.sched_getaffinity(0))
        xt = np.ascontiguousarray(x.permute(2, 1, 0).cpu().numpy())  # [gz][gy][gx] layout

        def port(m):
            sub = np.ascontiguousarray(lm.soa[:, :m].cpu().numpy())
            t0 = time.perf_counter()
            p = O.siddon_project(sub, xt, grid, center, size, 1.0, 1.0e6, prec="f32",
                                 nthreads=thr)
            O.siddon_backproject(sub, p.astype(np.float32), grid, center, size, 1.0, 1.0e6,
                                 prec="f32", nthreads=thr)
            return time.perf_counter() - t0

        # sizing run (also the warm-up), then the median of 3 runs of about a
        # quarter of the CPU budget each
        m = min(nl, 20000)
        dt = port(m)
        m2 = int(min(nl, max(m, m * args.cpu_seconds / 4 / max(dt, 1e-3))))
        runs = sorted(port(m2) for _ in range(3))
        dt = runs[1]
        cpu = {"value": m2 / dt / 1e6, "unit": "Mevents/s", "cores": thr, "kind": "port",
               "lscpu": lscpu_topology(),
               "sample": f"first {m2} LORs of subset 0, Siddon projection + ratio "
                         f"backprojection, FP32 port of siddon.cu.cc, {thr} threads; "
                         f"sizing run as warm-up, median of 3 runs"}
    print(json.dumps(_base(
        "Siddon OSEM Mevents/s (proj+backproj), non-TOF", value, "Mevents/s", args.steps,
        args.warmup, ms,
        {"workload": "siddon_c1", "image": list(grid), "voxel_mm": [3.0, 3.0, 3.4], "lors": n,
         "subsets": K, "step": "one OSEM iteration over all LORs",
         "path": "per subset (srfx_siddon_plan_osem_subset with counts): siddon_kernel<kProj> "
                 "over the plan's sorted LOR copy, sid_tile_scatter of c / p; mlem_update",
         "nnz_per_lor": round(nnz / nl, 1), "tile": list(plans[0].tile),
         "runs_per_lor": round(nruns / nl, 2), "chunks": nchunks,
         "untiled_backprojection_ms": round(ums, 4)},
        value_excl_plan_build=round(value_excl, 3), plan_build_ms=round(plan_s * 1e3, 2),
        plan_build_cold_ms=round(plan_times[0] * 1e3, 2),
        plan_build_amortised_over_iterations=C1_ITERATIONS,
        roofline={"bound": "hbm", "achieved": kern[dom]["achieved_GBps"], "peak": 8000.0,
                  "unit": "GB/s", "frac": kern[dom]["frac"],
                  "traffic": kern[dom].get("traffic_bytes_per_launch"),
                  "traffic_source": (traffic[dom]["source"] if traffic.get(dom) else None),
                  "kernel": kern[dom]["kernel"], "mean_launch_ms": kern[dom]["mean_launch_ms"],
                  "algorithmic_bytes_per_launch": byts,
                  "bytes_formula": "sum_i (28 + 4 + 4*nnz_i) + 4*N_vox"},
        kernels=kern, cpu_baseline=cpu)), flush=True)


# --------------------------------------------------------------------- C3 ---

def autograd_c3(args):
    import torch
    from srf_amd import autograd, physics, preprocess, synth
    from srf_amd.data import Image, split_from_groups
    from srf_amd.scanner import CylindricalScanner
    dev = torch.device("cuda:0")
    sc = CylindricalScanner()
    grid, center, size = (128, 128, 64), (0.0, 0.0, 0.0), (384.0, 384.0, 217.6)
    tb, ts2 = sc.tof_info()
    model = physics.SplitLoRsModel(4.0, False, tof_sigma2=ts2, tof_bin=tb)
    B, nev = 8, 2_000_000
    items = []
    for b in range(B):
        img = synth.lumpy_image(grid, center, size, seed=100 + b, device=dev)
        ev, _ = synth.tof_events_from_image(nev, img, center, size, seed=200 + b, scanner=sc,
                                            device=dev)
        data = split_from_groups(preprocess.recon_process(ev, sc.tof_res_ps))
        items.append((img, data))
    sens = torch.ones(grid, device=dev)
    # tile plans of every image's events, built once (like an OSEM subset's)
    t0 = time.time()
    plans = [model.make_plans(Image(img, center, size), data) for img, data in items]
    torch.cuda.synchronize()
    plan_s = time.time() - t0
    plan_bytes = sum(p.device_bytes() for ps in plans for p in ps.values() if p is not None)

    def grad_autograd():
        grads = []
        for (img, data), pl in zip(items, plans):
            x = img.clone().requires_grad_(True)
            p = autograd.project(Image(x, center, size), data, model, plans=pl)
            loss = autograd.poisson_nll(p.clamp_min(1e-7), x, sens)
            loss.backward()
            grads.append(x.grad)
        return grads

    def grad_plain():
        grads = []
        for (img, data), pl in zip(items, plans):
            with torch.no_grad():
                proj = model.projection(Image(img, center, size), data)
                bp = model.backprojection(proj, Image(img, center, size), plans=pl).data
                grads.append(sens - bp)
        return grads

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = _timer()
        a.record()
        for _ in range(args.steps):
            out = fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / args.steps, out

    torch.cuda.reset_peak_memory_stats()
    base_mem = torch.cuda.memory_allocated()
    ms_ag, g_ag = timed(grad_autograd)
    peak_ag = torch.cuda.max_memory_allocated() - base_mem
    ms_pl, g_pl = timed(grad_plain)
    err = max(float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) for a, b in zip(g_ag, g_pl))
    stored = sum(autograd.stored_list_bytes(model, Image(img, center, size), data)
                 for img, data in items)
    total = B * nev
    print(json.dumps(_base(
        "differentiable TOF projector Mevents/s (forward + backward)", total / (ms_ag / 1e3) / 1e6,
        "Mevents/s", args.steps, 1, ms_ag,
        {"workload": "autograd_c3", "images": B, "image": list(grid), "events_per_image": nev,
         "loss": "list-mode Poisson NLL", "step": "gradient of all 8 images"},
        grad_vs_plain={"autograd_ms": round(ms_ag, 3), "plain_ms": round(ms_pl, 3),
                       "overhead": round(ms_ag / ms_pl, 3), "max_rel_l2_grad_diff": err},
        memory={"recompute_peak_bytes": int(peak_ag), "stored_voxel_list_bytes": int(stored),
                "tile_plan_bytes": int(plan_bytes), "tile_plan_build_s": round(plan_s, 3),
                "note": "backward recomputes the weights, scattering through tile plans "
                        "(4 bytes per (event, tile) listing + 148 per event); a stored-list "
                        "backward keeps an int32 index and a float32 weight per non-zero "
                        "entry"})), flush=True)


# --------------------------------------------------------------------- C4 ---

def batched_c4(args):
    import torch
    from srf_amd import effmap, physics, preprocess, synth
    from srf_amd.batch import BatchedOSEM, serial_drivers, serial_iterate
    from srf_amd.data import Image, split_from_groups
    from srf_amd.scanner import CylindricalScanner
    dev = torch.device("cuda:0")
    sc = CylindricalScanner()
    tb, ts2 = sc.tof_info()
    model = physics.SplitLoRsModel(4.0, False, tof_sigma2=ts2, tof_bin=tb)
    kinds = [((96, 96, 48), (384.0, 384.0, 216.0)), ((128, 128, 64), (384.0, 384.0, 217.6)),
             ((200, 200, 100), (400.0, 400.0, 220.0))]
    center = (0.0, 0.0, 0.0)
    npairs = args.effmap_pairs or None
    emaps = {}
    for g, s in kinds:
        summed = effmap.compute_tor_effmap(sc, model, g, center, s, dev, max_pairs=npairs)
        emaps[g] = effmap.finalize(summed, g, center, s, sc.inner_radius, sc.axial_length)
    rng = np.random.default_rng(4)
    R, K, iters = 64, 8, 2
    recons, total = [], 0
    for r in range(R):
        g, s = kinds[int(rng.integers(0, 3))]
        n = int(np.exp(rng.uniform(np.log(1e5), np.log(5e6))))
        ph = synth.c2_phantom(seed=300 + r, scanner=sc)
        ev, _ = synth.tof_events(n, ph, seed=400 + r, scanner=sc, device=dev)
        data = split_from_groups(preprocess.recon_process(ev, sc.tof_res_ps))
        x0 = effmap.initial_image(g, center, s, sc.inner_radius, sc.axial_length, dev)
        recons.append({"image": Image(x0, center, s), "emap": Image(emaps[g], center, s),
                       "data": data, "x0": x0.clone()})
        total += n
    log(f"C4: {R} reconstructions, {total} events")

    def reset():
        for rc in recons:
            rc["image"].data.copy_(rc["x0"])

    # the plans are built once per event set and charged to the iterations
    # (all of C4's 2); timed after a warm-up build, as the iterations are
    builds = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.time()
        bat = BatchedOSEM(recons, model, nb_subsets=K)
        torch.cuda.synchronize()
        builds.append(time.time() - t0)
        if len(builds) < 2:
            bat.close()
    plan_b = builds[1]
    log(f"batch plan ({R * 3 * K} tile plans) in {plan_b:.2f}s (first build {builds[0]:.2f}s)")
    t0 = time.time()
    tune = bat.autotune(reps=2)
    log(f"projection schedules (lanes per event, events per item) tuned in "
        f"{time.time() - t0:.2f}s: {bat.schedule}")
    bat.capture()
    reset()
    bat.replay(1)  # warm-up
    reset()
    torch.cuda.synchronize()
    a, b = _timer()
    a.record()
    bat.replay(iters)
    b.record()
    torch.cuda.synchronize()
    ms_b = a.elapsed_time(b)
    out_b = [rc["image"].data.clone() for rc in recons]
    reset()
    t0 = time.time()
    drivers = serial_drivers(recons, model, K)
    torch.cuda.synchronize()
    plan_s = time.time() - t0
    log(f"serial tile plans in {plan_s:.2f}s")
    serial_iterate(recons, model, K, 1, drivers=drivers)  # warm-up
    reset()
    torch.cuda.synchronize()
    a, b = _timer()
    a.record()
    serial_iterate(recons, model, K, iters, drivers=drivers)
    b.record()
    torch.cuda.synchronize()
    ms_s = a.elapsed_time(b)
    err = max(float(torch.linalg.norm(x - rc["image"].data) / torch.linalg.norm(rc["image"].data))
              for x, rc in zip(out_b, recons))
    bat.close()
    tot_b = ms_b + plan_b * 1e3
    tot_s = ms_s + plan_s * 1e3
    print(json.dumps(_base(
        "batched TOF OSEM Mevents/s (proj+backproj), 64 reconstructions",
        total * iters / (tot_b / 1e3) / 1e6, "Mevents/s", iters, 1, tot_b / iters,
        {"workload": "batched_c4", "reconstructions": R, "events": total, "subsets": K,
         "iterations": iters, "image_sizes": [list(k[0]) for k in kinds],
         "step": "one OSEM iteration of all reconstructions",
         "projection_schedule": list(bat.schedule),
         "autotune_iteration_ms": {f"{l}x{e}": round(ms, 2) for (l, e), ms in tune.items()}},
        value_excl_plan_build=round(total * iters / (ms_b / 1e3) / 1e6, 3),
        plan_build_ms=round(plan_b * 1e3, 1), plan_build_cold_ms=round(builds[0] * 1e3, 1),
        plan_build_amortised_over_iterations=iters,
        batched_vs_serial={"batched_graph_ms": round(ms_b, 2), "serial_ms": round(ms_s, 2),
                           "speedup": round(ms_s / ms_b, 3),
                           "serial_plan_build_ms": round(plan_s * 1e3, 1),
                           "speedup_with_plan_builds": round(tot_s / tot_b, 3),
                           "serial_Mevents_per_s": round(total * iters / (ms_s / 1e3) / 1e6, 2),
                           "max_rel_l2_image_diff": err})), flush=True)


# --------------------------------------------------------------------- C5 ---

def large_c5(args):
    import torch
    from srf_amd import attenuation, effmap, physics, preprocess, synth
    from srf_amd.data import Image, split_from_groups
    from srf_amd.recon import OSEM
    from srf_amd.scanner import CylindricalScanner
    dev = torch.device("cuda:0")
    sc = CylindricalScanner(tof_res_ps=400.0)
    grid, center, size = (256, 256, 160), (0.0, 0.0, 0.0), (409.6, 409.6, 272.0)
    tb, ts2 = sc.tof_info()
    model = physics.SplitLoRsModel(4.0, False, tof_sigma2=ts2, tof_bin=tb)
    act, mu = synth.torso_image(grid, center, size, seed=5, device=dev)
    nev = args.events if args.events != 20_000_000 else 300_000_000
    t0 = time.time()
    groups = {a: [] for a in "xyz"}
    for c0 in range(0, nev, 25_000_000):
        m = min(25_000_000, nev - c0)
        ev, _ = synth.tof_events_from_image(m, act, center, size, seed=500 + c0 // 25_000_000,
                                            scanner=sc, device=dev)
        g = preprocess.recon_process(ev, sc.tof_res_ps)
        del ev
        for a in "xyz":
            groups[a].append(g[a])
    groups = {a: torch.cat(v, 1) for a, v in groups.items()}
    torch.cuda.synchronize()
    log(f"C5: {nev} events generated + preprocessed in {time.time() - t0:.1f}s")
    t0 = time.time()
    summed = attenuation.attenuated_tor_effmap(sc, model, Image(mu, center, size), grid, center,
                                               size, dev, max_pairs=args.effmap_pairs or None)
    E = effmap.finalize(summed, grid, center, size, sc.inner_radius, sc.axial_length)
    torch.cuda.synchronize()
    eff_s = time.time() - t0
    eff_stats = {"pairs": args.effmap_pairs or sc.nb_pairs(), "seconds": round(eff_s, 2),
                 "sum": float(summed.double().sum()),
                 "nonzero_voxels": int((summed > 0).sum()),
                 "inverse_map_nonzero_voxels": int((E > 0).sum())}
    log(f"attenuated efficiency map in {eff_s:.1f}s: {eff_stats}")
    x = effmap.initial_image(grid, center, size, sc.inner_radius, sc.axial_length, dev)
    K = 20
    t0 = time.time()
    groups, _ = preprocess.reorder_groups(groups, K)
    torch.cuda.synchronize()
    log(f"locality order inside the {K} subsets in {time.time() - t0:.1f}s")
    osem = OSEM(model, split_from_groups(groups), Image(E, center, size), nb_subsets=K,
                tiled=True)
    img = Image(x, center, size)
    # the plans are built once per event set and charged to C5's 2 iterations;
    # timed after a warm-up build, as in the other workloads
    builds = []
    for _ in range(2):
        osem.close()
        torch.cuda.synchronize()
        t0 = time.time()
        osem.plan(img)
        torch.cuda.synchronize()
        builds.append(time.time() - t0)
    plan_s = builds[1]
    listings = sum(p.info()[0] for ps in osem.plans for p in ps.values() if p is not None)
    log(f"tile plans in {plan_s:.2f}s, {listings / nev:.1f} listings per event, "
        f"{torch.cuda.memory_allocated() / 2**30:.1f} GiB allocated")
    torch.cuda.synchronize()
    a, b = _timer()
    a.record()
    osem.iterate(img, args.steps)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    assert torch.isfinite(x).all().item()
    ms_step = ms + plan_s * 1e3 / C5_ITERATIONS
    print(json.dumps(_base(
        "TOF OSEM Mevents/s (proj+backproj), large frame", nev / (ms_step / 1e3) / 1e6,
        "Mevents/s", args.steps, 0, ms_step,
        {"workload": "large_c5", "image": list(grid), "voxel_mm": [1.6, 1.6, 1.7],
         "events": nev, "subsets": K, "tof_fwhm_ps": 400, "attenuation": "mu 0.0096/mm body",
         "step": "one OSEM iteration over all events", "tiled": True, "sorted": True,
         "plan_build_s": round(plan_s, 3), "listings_per_event": round(listings / nev, 2),
         "peak_gib": round(torch.cuda.max_memory_allocated() / 2**30, 1),
         "effmap": eff_stats},
        value_excl_plan_build=round(nev / (ms / 1e3) / 1e6, 3),
        plan_build_ms=round(plan_s * 1e3, 1), plan_build_cold_ms=round(builds[0] * 1e3, 1),
        plan_build_amortised_over_iterations=C5_ITERATIONS)), flush=True)


# -------------------------------------------------------------------- PSF ---

def psf_c2(args):
    """Forward + backward PSF blur at 200x200x100 (srfx_psf_apply)."""
    import torch
    from srf_amd.data import Image
    from srf_amd.psf import gaussian_psf
    dev = torch.device("cuda:0")
    grid, size = (200, 200, 100), (400.0, 400.0, 220.0)
    t0 = time.time()
    psf = gaussian_psf(grid, size, 4.0, 4.4, dev)
    build_s = time.time() - t0
    x = Image(torch.rand(grid, device=dev) + 0.5, (0, 0, 0), size)
    f = torch.empty_like(x.data)
    b = torch.empty_like(x.data)
    for _ in range(max(args.warmup, 1)):
        psf.forward(x, out=f)
        psf.backward(x, out=b)
    torch.cuda.synchronize()
    a, e = _timer()
    a.record()
    for _ in range(args.steps):
        psf.forward(x, out=f)
        psf.backward(Image(f, x.center, size), out=b)
    e.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(e) / args.steps
    nvox = 200 * 200 * 100
    nnz = int(psf._fwd[2].numel())
    # per blur: image read + written, CSR (8 bytes per entry + row pointers), dense z matrix
    per_blur = 2 * 4 * nvox + 8 * nnz + 4 * (200 * 200 + 1) + 4 * 100 * 100
    gbs = 2 * per_blur / (ms / 1e3) / 1e9
    log(f"psf pair {ms * 1e3:.1f} us, {nnz / (200 * 200):.1f} transaxial entries per row")
    print(json.dumps(_base("PSF blur pair (forward + backward) per OSEM step", ms * 1e3, "us",
                           args.steps, args.warmup, ms,
                           {"workload": "psf_c2", "image": list(grid), "voxel_mm": [2.0, 2.0, 2.2],
                            "fwhm_mm": [4.0, 4.4], "xy_entries_per_row": nnz / (200 * 200),
                            "build_s": round(build_s, 2)},
                           higher_is_better=False,
                           roofline={"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0,
                                     "unit": "GB/s", "frac": round(gbs / 8000.0, 4),
                                     "traffic": None})), flush=True)


# --------------------------------------------------------------------- IO ---

def io_c2(args):
    """load_lors of a 20M-row float32 .npy file into device memory."""
    import tempfile
    import torch
    from srf_amd import io
    n = 20_000_000
    rows = np.random.default_rng(0).normal(0, 200, (n, 7)).astype(np.float32)
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        path = os.path.join(d, "lors.npy")
        np.save(path, rows)
        del rows
        io.load_lors(path, "cuda:0", dtype=torch.float32)  # warm the page cache
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(args.steps, 1)):
            t0 = time.perf_counter()
            out = io.load_lors(path, "cuda:0", dtype=torch.float32)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            del out
    s = float(np.median(ts))
    gbs = 28 * n / s / 1e9
    log(f"load_lors: {s * 1e3:.1f} ms for {n} rows, {gbs:.2f} GB/s")
    print(json.dumps(_base("list-mode .npy rows to device GB/s", gbs, "GB/s", len(ts), 1,
                           s * 1e3, {"workload": "io_c2", "rows": n, "bytes": 28 * n,
                                     "source": "page-cached file, pinned double buffer"})),
          flush=True)
