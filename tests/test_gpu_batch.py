"""Batched reconstructions in one persistent launch per subset
(srfx_tor_batch_*): equal to the same reconstructions run one after another,
with mixed image sizes (patch sizes 3, 4 and 5) and uneven event counts, and
stable under HIP-graph capture (several replays, each checked against eager
results, with eager work between replays)."""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

KW = 20.0
SHAPES = [((24, 24, 12), (384.0, 384.0, 216.0)),
          ((32, 32, 16), (384.0, 384.0, 216.0)),
          ((40, 40, 20), (384.0, 384.0, 216.0))]


def _recons(cuda, nrec=6, seed=0):
    from srf_amd import preprocess, synth
    from srf_amd.data import Image, split_from_groups
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner()
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nrec):
        shape, size = SHAPES[r % len(SHAPES)]
        n = int(rng.integers(2000, 20000))
        ev, _ = synth.tof_events(n, synth.c2_phantom(seed=10 + r, scanner=sc), seed=20 + r,
                                 scanner=sc)
        data = split_from_groups(preprocess.recon_process(ev.to(cuda), sc.tof_res_ps))
        x = torch.ones(shape, device=cuda)
        e = torch.from_numpy(rng.uniform(0.5, 1.5, shape).astype(np.float32)).to(cuda)
        out.append({"image": Image(x, (0, 0, 0), size), "emap": Image(e, (0, 0, 0), size),
                    "data": data})
    return out, sc


def _clone(recons):
    from srf_amd.data import Image
    return [{"image": Image(r["image"].data.clone(), r["image"].center, r["image"].size),
             "emap": r["emap"], "data": r["data"]} for r in recons]


def _model(sc):
    from srf_amd import physics
    tb, ts2 = sc.tof_info()
    return physics.SplitLoRsModel(KW, False, tof_sigma2=ts2, tof_bin=tb)


def test_batched_equals_serial(cuda):
    from srf_amd.batch import BatchedOSEM, serial_iterate
    recons, sc = _recons(cuda)
    model = _model(sc)
    a, b = _clone(recons), _clone(recons)
    bat = BatchedOSEM(a, model, nb_subsets=3)
    bat.iterate(2)
    serial_iterate(b, model, nb_subsets=3, nb_iterations=2)
    torch.cuda.synchronize()
    for ra, rb in zip(a, b):
        e = H.rel_l2(ra["image"].data.cpu().numpy(), rb["image"].data.cpu().numpy())
        assert e <= 1e-5, e
    bat.close()


def test_graph_replays_match_eager(cuda):
    from srf_amd.batch import BatchedOSEM
    recons, sc = _recons(cuda, nrec=4, seed=1)
    model = _model(sc)
    g_in = _clone(recons)
    e_in = _clone(recons)
    gb = BatchedOSEM(g_in, model, nb_subsets=2)
    eb = BatchedOSEM(e_in, model, nb_subsets=2)
    gb.capture()
    # capture records without running: reset the graph's images to the start
    for rg, r0 in zip(g_in, recons):
        rg["image"].data.copy_(r0["image"].data)
    for rep in range(3):
        gb.replay(1)
        eb.iterate(1)
        torch.cuda.synchronize()
        junk = torch.randn(1000, device=cuda).sum()  # eager work between replays
        torch.cuda.synchronize()
        for rg, re_ in zip(g_in, e_in):
            e = H.rel_l2(rg["image"].data.cpu().numpy(), re_["image"].data.cpu().numpy())
            assert e <= 1e-5, (rep, e)
        assert torch.isfinite(junk)
    gb.close()
    eb.close()


def test_projection_schedules_equal_default(cuda):
    """Every schedule of the persistent projection (lanes per event x events
    per work item) gives the default schedule's images: only the order of the
    lanes' partial sums of one event differs.  The autotuner keeps one of
    them and restores the images it iterated."""
    from srf_amd.batch import ITEM_EVENTS, LANES, BatchedOSEM
    recons, sc = _recons(cuda, nrec=4, seed=2)
    model = _model(sc)
    ref = _clone(recons)
    rb = BatchedOSEM(ref, model, nb_subsets=2)
    rb.iterate(1)
    torch.cuda.synchronize()
    for lanes in LANES:
        for events in ITEM_EVENTS:
            got = _clone(recons)
            b = BatchedOSEM(got, model, nb_subsets=2)
            b.set_schedule(lanes, events)
            b.iterate(1)
            torch.cuda.synchronize()
            for rg, rr in zip(got, ref):
                e = H.rel_l2(rg["image"].data.cpu().numpy(), rr["image"].data.cpu().numpy())
                assert e <= 1e-5, (lanes, events, e)
            b.close()
    with pytest.raises(RuntimeError):
        rb.set_schedule(2, 64)
    assert rb.schedule == (4, 64)
    tuned = _clone(recons)
    tb = BatchedOSEM(tuned, model, nb_subsets=2)
    table = tb.autotune(reps=1)
    assert set(table) == {(l, e) for l in LANES for e in ITEM_EVENTS}
    assert tb.schedule == min(table, key=table.get)
    for rt, r0 in zip(tuned, recons):
        assert torch.equal(rt["image"].data, r0["image"].data)
    tb.iterate(1)
    torch.cuda.synchronize()
    for rt, rr in zip(tuned, ref):
        assert H.rel_l2(rt["image"].data.cpu().numpy(), rr["image"].data.cpu().numpy()) <= 1e-5
    tb.close()
    rb.close()


# The batched workload's three image kinds (kernel_width 4: box 3) and the
# large frame's pitch (box 4), each a separate reconstruction of its own
# events, as the reference reconstructs one image per run
# (graph/reconstruction/local_reconstruction.py:44-51).
KINDS_KW4 = [((96, 96, 48), (384.0, 384.0, 216.0), 150.0),
             ((128, 128, 64), (384.0, 384.0, 217.6), 150.0),
             ((200, 200, 100), (400.0, 400.0, 220.0), 150.0),
             ((64, 64, 40), (102.4, 102.4, 68.0), 45.0)]


def _recons_kw4(cuda, seed=7):
    from srf_amd import preprocess, synth
    from srf_amd.data import Image, split_from_groups
    from srf_amd.scanner import CylindricalScanner
    sc = CylindricalScanner()
    rng = np.random.default_rng(seed)
    out = []
    for r, (shape, size, radius) in enumerate(KINDS_KW4 + KINDS_KW4[:1]):
        n = int(rng.integers(6000, 16000))
        ph = synth.Phantom(radius, 0.45 * size[2] if radius < 100 else sc.axial_length / 2,
                           [((0.3 * radius, -0.2 * radius, 0.0), 0.15 * radius, 4.0)])
        ev, _ = synth.tof_events(n, ph, seed=30 + r, scanner=sc, device=cuda)
        groups = preprocess.recon_process(ev, sc.tof_res_ps)
        data = split_from_groups(groups)
        x = torch.ones(shape, device=cuda)
        e = torch.from_numpy(rng.uniform(0.5, 1.5, shape).astype(np.float32)).to(cuda)
        out.append({"image": Image(x, (0, 0, 0), size), "emap": Image(e, (0, 0, 0), size),
                    "data": data, "groups": {a: groups[a].cpu().numpy() for a in "xyz"}})
    return out, sc


def test_batched_graph_vs_oracle(cuda):
    """Mixed reconstructions at the batched workload's pitches and the large
    frame's, kernel_width 4, 2 iterations x 8 subsets, run as one captured HIP
    graph per iteration: each image equals the oracle's OSEM of that
    reconstruction alone (rel L2 <= 1e-4)."""
    from oracle import recon as R
    from srf_amd import _lib
    from srf_amd.batch import BatchedOSEM
    recons, sc = _recons_kw4(cuda)
    tb, ts2 = sc.tof_info()
    boxes = {_lib.tor_patch_info(*_frame(rc), 4.0)[1] for rc in recons}
    assert boxes == {3, 4}
    model = _model_kw(sc, 4.0)
    start = [rc["image"].data.clone() for rc in recons]
    bat = BatchedOSEM(recons, model, nb_subsets=8)
    bat.capture()
    for rc, x0 in zip(recons, start):  # capture records without running
        rc["image"].data.copy_(x0)
    bat.replay(2)
    torch.cuda.synchronize()
    for r, (rc, x0) in enumerate(zip(recons, start)):
        ref = R.osem(rc["groups"], x0.cpu().numpy(), rc["emap"].data.cpu().numpy().astype(
            np.float64), (0, 0, 0), rc["image"].size, 4.0, False, tb, ts2, 2, 8)
        e = H.rel_l2(rc["image"].data.cpu().numpy(), ref)
        n = sum(rc["groups"][a].shape[1] for a in "xyz")
        print(f"batched recon {r} {tuple(rc['image'].data.shape)}: {n} events, "
              f"rel_l2 vs oracle {e:.2e}")
        assert e <= 1e-4
    bat.close()


def _frame(rc):
    from srf_amd import physics
    x = rc["image"].data
    g, c, s, _ = physics.group_frame(x.shape, (0, 0, 0), rc["image"].size, (1, 1, 1), "x")
    return g, s


def _model_kw(sc, kw):
    from srf_amd import physics
    tb, ts2 = sc.tof_info()
    return physics.SplitLoRsModel(kw, False, tof_sigma2=ts2, tof_bin=tb)
