"""Pins the oracle's OSEM driver (oracle/recon.py, the semantics of
ReconStep + mlem_update, recon_step.py:33-49, over contiguous subsets,
worker.py:79-85) with identities that do not reuse its code:

* list-mode count identity: after one update x' = x E (A^T (1 / A x)) with
  E = 1 / s on the support of s = A^T 1 (the un-normalised sensitivity of the
  subset's events), sum_j s_j x'_j = #{i : (A x)_i > 1e-7} -- every event
  whose projection passes the threshold contributes exactly 1;
* monotone Poisson log-likelihood for K = 1 (MLEM): with E = 1 / s the
  iteration is the EM algorithm for list-mode data, whose log-likelihood
  L(x) = sum_i log (A x)_i - sum_j s_j x_j never decreases;
* the same two for the Siddon driver (complete_lors.py:59-98);
* a dropped term (no effmap factor, or the 1e-7 threshold ignored) breaks the
  count identity.
"""
import numpy as np
import pytest

from oracle import lib as O
from oracle import recon as R
from tests import helpers as H

WORLD = ((20, 18, 12), (0.5, -1.0, 0.25), (40.0, 36.0, 26.4))  # (nx, ny, nz) world, 2x2x2.2 mm
TOF = (H.TOF_BIN, H.TOF_SIGMA2)


def _groups(n, seed):
    """Events of the three axis groups in their op frames (rows as
    on_tor_lors.py produces them), drawn inside the world box."""
    grid, center, size = WORLD
    out = {}
    for k, a in enumerate("xyz"):
        dummy = np.zeros(grid, np.float32)
        _, g, c, s = R.group_frame(dummy, center, size, a)
        out[a] = H.tor_lors(n if a != "z" else n // 4, seed=seed + k, grid=tuple(g),
                            center=tuple(c), size=tuple(s))
    return out


def _sensitivity(groups):
    grid, center, size = WORLD
    ones = {a: np.ones(groups[a].shape[1], np.float32) for a in groups}
    return R.backproject_split(groups, ones, grid, center, size, H.KW, False, *TOF)


def _proj(groups, x):
    grid, center, size = WORLD
    p = R.project_split(groups, x.astype(np.float32), center, size, H.KW, False, *TOF)
    return np.concatenate([p[a] for a in "xyz"])


def test_count_identity_one_update():
    grid, center, size = WORLD
    groups = _groups(300, seed=5)
    s = _sensitivity(groups)
    E = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    x0 = H.image(grid[::-1], seed=2).reshape(grid)  # positive (nx, ny, nz)
    x1 = R.osem(groups, x0, E, center, size, H.KW, False, *TOF, nb_iterations=1, nb_subsets=1)
    npos = int(np.sum(_proj(groups, x0) > 1e-7))
    assert npos > 300
    assert float(np.sum(s * x1)) == pytest.approx(npos, rel=1e-6)
    # a dropped effmap factor breaks it
    x_bad = R.osem(groups, x0, np.ones_like(E), center, size, H.KW, False, *TOF,
                   nb_iterations=1, nb_subsets=1)
    assert abs(float(np.sum(s * x_bad)) - npos) > 1e-3 * npos


def test_count_identity_threshold():
    """Events whose projection is at most 1e-7 are left out of the ratio
    (tor.cu.cc:203-206): the identity counts only the others."""
    grid, center, size = WORLD
    groups = _groups(200, seed=9)
    s = _sensitivity(groups)
    E = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    x0 = np.full(grid, 1e-9)  # every projection falls below the threshold
    x0[grid[0] // 2] = 1.0   # except events through one slab of the image
    x1 = R.osem(groups, x0, E, center, size, H.KW, False, *TOF, nb_iterations=1, nb_subsets=1)
    p = _proj(groups, x0)
    npos = int(np.sum(p > 1e-7))
    assert 0 < npos < p.size
    assert float(np.sum(s * x1)) == pytest.approx(npos, rel=1e-6)


def test_mlem_loglikelihood_monotone():
    grid, center, size = WORLD
    groups = _groups(250, seed=13)
    s = _sensitivity(groups)
    E = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    x = np.where(s > 0, 1.0, 0.0)

    def loglik(x):
        p = _proj(groups, x)
        keep = p > 1e-7
        return float(np.sum(np.log(p[keep])) - np.sum(s * x))

    ll = [loglik(x)]
    for _ in range(4):
        x = R.osem(groups, x, E, center, size, H.KW, False, *TOF, nb_iterations=1, nb_subsets=1)
        ll.append(loglik(x))
    assert all(b >= a - 1e-9 * abs(a) for a, b in zip(ll, ll[1:])), ll
    assert ll[-1] > ll[0]


SID = (H.SID_GRID, H.SID_CENTER, H.SID_SIZE)


def _sid_world_sensitivity(lors):
    grid, center, size = SID
    b = O.siddon_backproject(lors, np.ones(lors.shape[1], np.float32), grid, center, size, 1.0,
                             1.0e6, prec="f64")
    return np.transpose(b, (2, 1, 0))  # [gz][gy][gx] -> world (nx, ny, nz)


def test_siddon_count_identity_and_loglik():
    grid, center, size = SID
    lors = H.siddon_lors(400, seed=3, special=False)
    w = np.random.default_rng(4).integers(1, 5, lors.shape[1]).astype(np.float32)  # counts
    s = _sid_world_sensitivity(lors)
    E = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    x0 = np.where(s > 0, 1.0, 0.0)
    x1 = R.siddon_osem(lors, x0, E, center, size, 1.0, 1.0e6, nb_iterations=1, nb_subsets=1,
                       weights=w)
    p = O.siddon_project(lors, np.transpose(x0, (2, 1, 0)).astype(np.float32), grid, center,
                         size, 1.0, 1.0e6, prec="f64")
    # weighted identity: every event with p > 0 contributes its count
    assert float(np.sum(s * x1)) == pytest.approx(float(w[p > 0].sum()), rel=1e-6)

    def loglik(x):
        q = O.siddon_project(lors, np.transpose(x, (2, 1, 0)).astype(np.float32), grid, center,
                             size, 1.0, 1.0e6, prec="f64")
        keep = q > 0
        return float(np.sum(w[keep] * np.log(q[keep])) - np.sum(s * x))

    x, ll = x0, [loglik(x0)]
    for _ in range(3):
        x = R.siddon_osem(lors, x, E, center, size, 1.0, 1.0e6, nb_iterations=1, nb_subsets=1,
                          weights=w)
        ll.append(loglik(x))
    assert all(b >= a - 1e-6 * abs(a) for a, b in zip(ll, ll[1:])), ll
    assert ll[-1] > ll[0]
