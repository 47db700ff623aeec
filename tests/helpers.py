"""Seeded synthetic inputs for parity tests, directly in an op's frame."""
import numpy as np

# A small TOR frame: x, y in-slice, z = slice axis (the group's dominant axis).
TOR_GRID = (40, 36, 24)          # gx, gy, gz
TOR_CENTER = (1.0, -2.0, 0.5)
TOR_SIZE = (80.0, 72.0, 52.8)    # 2.0 x 2.0 x 2.2 mm voxels
TOF_BIN = 6.0
TOF_SIGMA2 = 1139.77             # 530 ps FWHM, srfapp_new.py:244-251
KW = 4.0


def tor_lors(n, seed=0, grid=TOR_GRID, center=TOR_CENTER, size=TOR_SIZE, full=False):
    """[10, n] float32 LORs in the op frame: end points below and above the
    box along z (|dz| dominant, dz >= 0 as after SwapPointsOrder), a TOF
    centre on the segment and a sigma2 factor.  A fraction of the events
    end inside the box or miss it in x/y."""
    rng = np.random.default_rng(seed)
    c = np.array(center, np.float64)
    s = np.array(size, np.float64)
    lo, hi = c - s / 2, c + s / 2
    p1 = np.empty((n, 3))
    p2 = np.empty((n, 3))
    p1[:, 0] = rng.uniform(lo[0] - 10, hi[0] + 10, n)
    p1[:, 1] = rng.uniform(lo[1] - 10, hi[1] + 10, n)
    p2[:, 0] = p1[:, 0] + rng.uniform(-25, 25, n)
    p2[:, 1] = p1[:, 1] + rng.uniform(-25, 25, n)
    p1[:, 2] = rng.uniform(lo[2] - 30, lo[2] + 5, n)
    p2[:, 2] = rng.uniform(hi[2] - 5, hi[2] + 30, n)
    # keep z dominant (the group's slice axis)
    dz = p2[:, 2] - p1[:, 2]
    for k in (0, 1):
        d = p2[:, k] - p1[:, k]
        bad = np.abs(d) > dz
        p2[bad, k] = p1[bad, k] + np.sign(d[bad]) * dz[bad] * rng.uniform(0, 1, bad.sum())
    r = rng.uniform(0.15, 0.85, n)
    pc = p1 + r[:, None] * (p2 - p1)
    if full:
        f = rng.uniform(0.3, 1.7, n)
    else:
        f = rng.uniform(0.5, 1.5, n)
    out = np.vstack([p1.T, p2.T, pc.T, f[None, :]]).astype(np.float32)
    return out


def tor_map_lors(n, seed=0, **kw):
    l10 = tor_lors(n, seed=seed, **kw)
    return np.vstack([l10[0:6], l10[9:10]])


def image(grid, seed=0, positive=True):
    rng = np.random.default_rng(seed + 1000)
    gx, gy, gz = grid
    img = rng.uniform(0.2, 2.0, (gz, gy, gx)) if positive else rng.normal(0, 1, (gz, gy, gx))
    return img.astype(np.float32)


SID_GRID = (24, 20, 16)
SID_CENTER = (0.5, -1.0, 0.25)
SID_SIZE = (72.0, 60.0, 54.4)


def siddon_lors(n, seed=0, grid=SID_GRID, center=SID_CENTER, size=SID_SIZE, special=True):
    """[7, n] float32 rays between points on a sphere around the box, with a
    TOF offset; plus axis-aligned rays and rays lying on voxel planes."""
    rng = np.random.default_rng(seed)
    c = np.array(center, np.float64)
    s = np.array(size, np.float64)
    R = 0.75 * np.linalg.norm(s)
    def sphere(m):
        v = rng.normal(size=(m, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)
    p1 = c + R * sphere(n) + rng.uniform(-8, 8, (n, 3))
    p2 = c + R * sphere(n) + rng.uniform(-8, 8, (n, 3))
    L = np.linalg.norm(p2 - p1, axis=1)
    tof = rng.normal(0, 0.15, n) * L
    rows = np.vstack([p1.T, p2.T, tof[None, :]])
    if special:
        extra = []
        v = s / np.array(grid)
        for k in range(3):
            for j in range(6):
                a = c - s / 2 + v * (rng.integers(0, np.array(grid)) + rng.choice([0.5, 0.0, 0.25]))
                b = a.copy()
                a[k] = c[k] - s[k]
                b[k] = c[k] + s[k]
                if j % 2:
                    a, b = b, a
                extra.append(np.concatenate([a, b, [rng.normal(0, 5)]]))
        rows = np.hstack([rows, np.array(extra).T])
    return rows.astype(np.float32)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
