"""LOR preprocessing (on_tor_lors.py): the oracle against the reference's own
known answers, closed forms of the sigma2 factor and the TOF cut, and the
product's torch implementation against the oracle bit for bit."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import preprocess as OP

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "on_tor_lors_known_answers.json")))


def test_known_answers_direction_partition():
    lors = np.array(GOLD["lors"], np.float32)
    np.testing.assert_array_equal(OP.direction(lors[0:1]), GOLD["direction_row0"])
    d = np.eye(3, dtype=np.float32)
    assert OP.dominent_by("z", d[:, 0], d[:, 1], d[:, 2]).tolist() == GOLD[
        "dominent_by_z_of_unit_axes"][0]
    part, _ = OP.partition("x", lors)
    np.testing.assert_array_equal(part, GOLD["partition_x"])


def test_known_answers_map_process():
    lors = np.array(GOLD["lors"], np.float32)
    out = OP.map_process(lors)
    for a in "xyz":
        np.testing.assert_array_equal(out[a][:, 0:6], GOLD["map_process_first6"][a])


def test_product_map_process_matches_known_answers():
    from srf_amd import preprocess as P
    lors = torch.tensor(GOLD["lors"], dtype=torch.float32)
    out = P.map_process(lors)
    for a in "xyz":
        np.testing.assert_array_equal(out[a].t()[:, 0:6].numpy(), GOLD["map_process_first6"][a])


def test_sigma2_factor_closed_form():
    # transaxial LOR through the axis: dsq = (2R)^2, Rsq = 2R^2, |d|^2 = 4R^2
    # -> f = 16R^4 / 2R^2 / 4R^2 / 2 = 1; an axial LOR has f = 0
    R = 400.0
    lors = np.array([[R, 0, 0, -R, 0, 0], [3.0, 4.0, -10.0, 3.0, 4.0, 10.0]])
    f = OP.sigma2_factor(lors)
    np.testing.assert_allclose(f, [1.0, 0.0], atol=1e-15)


def test_cut_closed_form():
    """530 ps: limit = 530*0.15/2.35482005*3 = 101.2816... mm.  A cut LOR
    keeps 2*limit around centre = p1 + (0.5 - tof/L) d; one whose TOF centre is
    within the limit of p1 is left whole (the cut is asymmetric)."""
    limit = 530 * 0.15 / 2.35482005 * 3
    rows = np.array([[-400.0, 10.0, 5.0, 400.0, -10.0, -5.0, 30.0, 0.7],
                     [-400.0, 0.0, 0.0, 400.0, 0.0, 0.0, 350.0, 0.9]])
    out = OP.cut_lors(530.0, rows)
    d = rows[0, 3:6] - rows[0, 0:3]
    L = np.linalg.norm(d)
    c = rows[0, 0:3] + (0.5 - 30.0 / L) * d
    np.testing.assert_allclose(out[0, 6:9], c, rtol=1e-14)
    np.testing.assert_allclose(np.linalg.norm(out[0, 3:6] - out[0, 0:3]), 2 * limit, rtol=1e-12)
    np.testing.assert_allclose(out[1, 0:6], rows[1, 0:6])  # centre 50 mm from p1: uncut
    assert out[0, 9] == 0.7
    # no cut at all when tof_res >= 10000
    np.testing.assert_array_equal(OP.cut_lors(1.0e4, rows)[:, 0:6], rows[:, 0:6])


def test_swap_negates_tof_only_with_extra_columns():
    rows = np.array([[5.0, 0, 0, -5.0, 0, 0, 2.0, 0.3]])
    g, _ = OP.swap_points_order("x", rows)
    np.testing.assert_array_equal(g, [[-5.0, 0, 0, 5.0, 0, 0, -2.0, 0.3]])
    g7, _ = OP.swap_points_order("x", rows[:, :7])
    np.testing.assert_array_equal(g7, [[-5.0, 0, 0, 5.0, 0, 0, 2.0]])


def _events(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 2 * np.pi, n)
    b = a + rng.uniform(0.5, 2 * np.pi - 0.5, n)
    z1, z2 = rng.uniform(-54, 54, n), rng.uniform(-54, 54, n)
    p1 = np.column_stack([400 * np.cos(a), 400 * np.sin(a), z1])
    p2 = np.column_stack([400 * np.cos(b), 400 * np.sin(b), z2])
    L = np.linalg.norm(p2 - p1, axis=1)
    tof = rng.uniform(-0.4, 0.4, n) * L
    rows = np.column_stack([p1, p2, tof])
    # a few axial (z-group) rows and exact ties between |dx| and |dy|
    rows[0] = [10, 20, -60, 12, 21, 60, 3.0]
    rows[1] = [-300, -300, 0, 300, 300, 5, 1.0]
    rows[2] = [300, -300, 0, -300, 300, 5, -1.0]
    return rows


@pytest.mark.parametrize("tof_res", [530.0, 1.0e4])
def test_product_recon_process_bit_exact(tof_res):
    from srf_amd import preprocess as P
    rows = _events(5000, seed=1)
    ref = OP.recon_process(rows, tof_res)
    got, order = P.recon_process(torch.from_numpy(rows), tof_res, return_order=True)
    ref_order = OP.group_order(np.hstack([rows, OP.sigma2_factor(rows)[:, None]]))
    for a in "xyz":
        np.testing.assert_array_equal(got[a].t().numpy(), ref[a].astype(np.float32))
        np.testing.assert_array_equal(order[a].numpy(), ref_order[a])
    assert sum(ref[a].shape[0] for a in "xyz") == rows.shape[0]


def test_product_map_process_bit_exact():
    from srf_amd import preprocess as P
    rows = _events(3000, seed=2)[:, :6]
    ref = OP.map_process(rows)
    got = P.map_process(torch.from_numpy(rows))
    for a in "xyz":
        np.testing.assert_array_equal(got[a].t().numpy(), ref[a].astype(np.float32))


def test_tof_info():
    tb, ts2 = OP.tof_info(530.0, 40.0)
    assert tb == pytest.approx(6.0)
    assert np.sqrt(ts2) == pytest.approx(530 * 0.15 / 2.35482005)


def test_locality_order_keeps_subsets():
    """Reordering for cache locality permutes rows only inside each OSEM
    subset [k N / K, (k+1) N / K), so subset membership (worker.py:79-85) and
    therefore the OSEM result are unchanged."""
    from srf_amd import preprocess as P
    rows = _events(4000, seed=5)
    g = P.recon_process(torch.from_numpy(rows), 530.0)
    for K in (1, 3, 7):
        out, orders = P.reorder_groups(g, K)
        for a in "xyz":
            n = g[a].shape[1]
            o = orders[a].numpy()
            assert sorted(o.tolist()) == list(range(n))
            for k in range(K):
                lo, hi = k * n // K, (k + 1) * n // K
                assert set(o[lo:hi].tolist()) == set(range(lo, hi))
            np.testing.assert_array_equal(out[a].numpy(), g[a].numpy()[:, o])
