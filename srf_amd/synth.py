"""Seeded synthetic phantoms and list-mode data for the benchmark workloads.

Phantom parameters come from numpy.random.Generator(PCG64(seed)); event
streams from a seeded torch.Generator on the requested device (so they depend
on seed and device type).  Every parameter that shapes the data is returned in
the ``meta`` dict.

Workloads
  siddon_c1  non-TOF Siddon, 128x128x64 image (3.0x3.0x3.4 mm), 1M random
             crystal pairs (transaxial separation >= 90 deg), Poisson counts
             of the Siddon projection of a cylinder + 3 hot spheres.
  tof_c2     TOF list mode, 200x200x100 image (2.0x2.0x2.2 mm), cylinder +
             6 spheres, 20M events from emission points drawn from the
             phantom, isotropic directions, crystal-snapped end points and a
             Gaussian TOF offset (530 ps FWHM).
"""
import numpy as np

from .scanner import CylindricalScanner

GAUSSIAN_FACTOR = 2.35482005


def tof_sigma_mm(tof_res_ps):
    return tof_res_ps * 0.15 / GAUSSIAN_FACTOR


class Phantom:
    """Cylinder of value ``background`` (radius, half height) plus
    non-overlapping spheres fully inside it whose value replaces the
    background."""

    def __init__(self, radius, half_height, spheres, background=1.0):
        self.radius = float(radius)
        self.half_height = float(half_height)
        self.spheres = [(np.asarray(c, np.float64), float(r), float(a)) for c, r, a in spheres]
        self.background = float(background)

    def value(self, pts):
        pts = np.asarray(pts, np.float64)
        r2 = pts[..., 0] ** 2 + pts[..., 1] ** 2
        inside = (r2 <= self.radius ** 2) & (np.abs(pts[..., 2]) <= self.half_height)
        v = np.where(inside, self.background, 0.0)
        for c, r, a in self.spheres:
            d2 = np.sum((pts - c) ** 2, axis=-1)
            v = np.where(d2 <= r * r, a, v)
        return v

    def sample(self, gen, n, device):
        """n emission points (torch float64 [n, 3] on ``device``) distributed
        with density proportional to value(), drawn with torch generator gen."""
        import torch
        vcyl = np.pi * self.radius ** 2 * 2 * self.half_height * self.background
        w = [vcyl] + [4 / 3 * np.pi * r ** 3 * (a - self.background) for _, r, a in self.spheres]
        w = torch.tensor(np.asarray(w) / np.sum(w), dtype=torch.float64, device=device)
        comp = torch.multinomial(w, n, replacement=True, generator=gen)
        u = torch.rand(n, 3, dtype=torch.float64, device=device, generator=gen)
        rr = self.radius * torch.sqrt(u[:, 0])
        ph = 2 * np.pi * u[:, 1]
        out = torch.stack([rr * torch.cos(ph), rr * torch.sin(ph),
                           (2 * u[:, 2] - 1) * self.half_height], 1)
        v = torch.randn(n, 3, dtype=torch.float64, device=device, generator=gen)
        v = v / torch.linalg.norm(v, dim=1, keepdim=True)
        rad = torch.pow(torch.rand(n, dtype=torch.float64, device=device, generator=gen), 1 / 3)
        for i, (c, r, _) in enumerate(self.spheres, start=1):
            m = comp == i
            ct = torch.tensor(c, dtype=torch.float64, device=device)
            out[m] = ct + v[m] * (r * rad[m])[:, None]
        return out

    def voxelize(self, grid, center, size, sub=2):
        """Image (nx, ny, nz) of the mean value over sub^3 points per voxel."""
        grid = np.asarray(grid)
        center = np.asarray(center, np.float64)
        size = np.asarray(size, np.float64)
        vox = size / grid
        img = np.zeros(tuple(grid), np.float64)
        offs = (np.arange(sub) + 0.5) / sub - 0.5
        axes = [center[k] - size[k] / 2 + (np.arange(grid[k]) + 0.5) * vox[k] for k in range(3)]
        for ox in offs:
            for oy in offs:
                for oz in offs:
                    X, Y, Z = np.meshgrid(axes[0] + ox * vox[0], axes[1] + oy * vox[1],
                                          axes[2] + oz * vox[2], indexing="ij")
                    img += self.value(np.stack([X, Y, Z], -1))
        return (img / sub ** 3).astype(np.float32)


def _place_spheres(rng, radii, acts, cyl_r, half_h):
    spheres = []
    for r, a in zip(radii, acts):
        for _ in range(10000):
            rho = (cyl_r - r - 2) * np.sqrt(rng.uniform(0, 1))
            ph = rng.uniform(0, 2 * np.pi)
            z = rng.uniform(-half_h + r + 1, half_h - r - 1)
            c = np.array([rho * np.cos(ph), rho * np.sin(ph), z])
            if all(np.linalg.norm(c - c2) > r + r2 + 1 for c2, r2, _ in spheres):
                spheres.append((c, r, a))
                break
        else:
            raise RuntimeError("could not place spheres")
    return spheres


def c2_phantom(seed=2, scanner=CylindricalScanner()):
    rng = np.random.default_rng(seed)
    half_h = scanner.axial_length / 2
    radii = rng.uniform(5, 20, 6)
    acts = rng.uniform(2, 8, 6)
    return Phantom(150.0, half_h, _place_spheres(rng, radii, acts, 150.0, half_h))


def tof_events(n, phantom, seed=2, scanner=CylindricalScanner(), device="cpu",
               batch=8_000_000, gen=None):
    """List-mode TOF events, torch float64 [n, 7] on ``device``:
    x1,y1,z1,x2,y2,z2,tof.

    Emission point from the phantom density, direction uniform on the sphere
    (restricted to the |u_z| band that can reach both ends of the scanner, which
    leaves the accepted distribution unchanged), line intersected with the
    detector cylinder, kept when both ends are within the axial extent, end
    points snapped to crystal centres, tof = L/2 - |e' - p1| + N(0, sigma) with
    e' the emission point projected on the snapped LOR (the sign used by
    CutLoRs: positive tof = closer to p1).  The stream is seeded with a torch
    generator on ``device``, so the events depend on seed and device type."""
    import torch
    if gen is None:
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
    R = scanner.radius
    half = scanner.axial_length / 2
    sigma = tof_sigma_mm(scanner.tof_res_ps)
    rmax = phantom.radius
    chord = 2 * np.sqrt(R * R - rmax * rmax)
    uz_max = min(1.0, 1.05 * (2 * half) / np.sqrt(chord ** 2 + (2 * half) ** 2))
    nc = scanner.nb_crystals
    step = 2 * np.pi / nc
    z0 = float(scanner.ring_z()[0])
    pitch = scanner.ring_pitch + scanner.gap
    ang = torch.tensor(scanner.crystal_angles(), dtype=torch.float64, device=device)
    rz = torch.tensor(scanner.ring_z(), dtype=torch.float64, device=device)

    def snap(p):
        a = torch.atan2(p[:, 1], p[:, 0])
        cr = torch.remainder(torch.round(a / step).long(), nc)
        ring = torch.clamp(torch.round((p[:, 2] - z0) / pitch).long(), 0, scanner.nb_rings - 1)
        return torch.stack([R * torch.cos(ang[cr]), R * torch.sin(ang[cr]), rz[ring]], 1)

    out = []
    have = 0
    while have < n:
        m = min(batch, 3 * (n - have) + 1000)
        e = phantom.sample(gen, m, device)
        r = torch.rand(m, 2, dtype=torch.float64, device=device, generator=gen)
        uz = (2 * r[:, 0] - 1) * uz_max
        ph = 2 * np.pi * r[:, 1]
        s = torch.sqrt(1 - uz * uz)
        u = torch.stack([s * torch.cos(ph), s * torch.sin(ph), uz], 1)
        a = u[:, 0] ** 2 + u[:, 1] ** 2
        b = 2 * (e[:, 0] * u[:, 0] + e[:, 1] * u[:, 1])
        c = e[:, 0] ** 2 + e[:, 1] ** 2 - R * R
        disc = torch.sqrt(b * b - 4 * a * c)
        p1 = e + ((-b - disc) / (2 * a))[:, None] * u
        p2 = e + ((-b + disc) / (2 * a))[:, None] * u
        keep = (p1[:, 2].abs() <= half) & (p2[:, 2].abs() <= half)
        e, p1, p2 = e[keep], p1[keep], p2[keep]
        q1, q2 = snap(p1), snap(p2)
        d = q2 - q1
        L = torch.linalg.norm(d, dim=1)
        ok = L > 1.0
        e, q1, d, L = e[ok], q1[ok], d[ok], L[ok]
        t = torch.sum((e - q1) * d, dim=1) / L
        noise = torch.randn(L.shape[0], dtype=torch.float64, device=device, generator=gen)
        tof = L / 2 - t + sigma * noise
        rows = torch.cat([q1, q1 + d, tof[:, None]], 1)
        out.append(rows[: n - have])
        have += out[-1].shape[0]
    return torch.cat(out), {"sigma_tof_mm": sigma, "uz_max": uz_max, "seed": seed}


def c2(n_events=20_000_000, seed=2, scanner=CylindricalScanner(), device="cpu"):
    """TOF-TOR workload: image geometry, phantom and raw events (torch)."""
    ph = c2_phantom(seed, scanner)
    ev, meta = tof_events(n_events, ph, seed=seed + 100, scanner=scanner, device=device)
    tof_bin, tof_sigma2 = scanner.tof_info()
    meta.update(dict(grid=(200, 200, 100), voxel=(2.0, 2.0, 2.2), center=(0.0, 0.0, 0.0),
                     size=(400.0, 400.0, 220.0), kernel_width=4.0, geo_sigma2_flag=False,
                     tof_bin=tof_bin, tof_sigma2=tof_sigma2, n_events=n_events,
                     spheres=[(c.tolist(), r, a) for c, r, a in ph.spheres]))
    return ph, ev, meta


def c1_phantom(seed=1):
    rng = np.random.default_rng(seed)
    half_h = 64 * 3.4 / 2
    return Phantom(150.0, half_h, _place_spheres(rng, [10.0, 15.0, 20.0], [4.0, 4.0, 4.0],
                                                 150.0, half_h))


def c1_pairs(n, seed=1, scanner=CylindricalScanner(), min_sep=96):
    """n random unordered crystal pairs with transaxial separation >= min_sep
    crystals (90 deg at 384 per ring) and any ring difference: (n, 6) float32."""
    rng = np.random.default_rng(seed)
    nc = scanner.nb_crystals
    c1 = rng.integers(0, nc, n)
    sep = rng.integers(min_sep, nc - min_sep + 1, n)
    c2 = (c1 + sep) % nc
    r1 = rng.integers(0, scanner.nb_rings, n)
    r2 = rng.integers(0, scanner.nb_rings, n)
    p1 = scanner.crystal_xyz(r1, c1)
    p2 = scanner.crystal_xyz(r2, c2)
    return np.hstack([p1, p2]).astype(np.float32)


def tof_events_from_image(n, image, center, size, seed=2, scanner=CylindricalScanner(),
                          tof_res_ps=None, device="cpu", batch=8_000_000):
    """List-mode TOF events like tof_events, with emission points drawn from a
    voxelized activity image (torch (nx, ny, nz), non-negative): a voxel by its
    value, then a uniform point inside it."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    img = image.to(device=device, dtype=torch.float64).reshape(-1)
    grid = list(image.shape)
    vox = torch.tensor([size[k] / grid[k] for k in range(3)], dtype=torch.float64, device=device)
    lo = torch.tensor([center[k] - size[k] / 2 for k in range(3)], dtype=torch.float64,
                      device=device)
    strides = [grid[1] * grid[2], grid[2], 1]

    class _ImagePhantom:
        radius = float(min(np.hypot(size[0], size[1]) / 2, 0.99 * scanner.radius))

        @staticmethod
        def sample(g, m, dev):
            idx = torch.multinomial(img, m, replacement=True, generator=g)
            ijk = torch.stack([idx // strides[0], (idx // strides[1]) % grid[1],
                               idx % grid[2]], 1).to(torch.float64)
            u = torch.rand(m, 3, dtype=torch.float64, device=dev, generator=g)
            return lo + (ijk + u) * vox

    sc = scanner
    if tof_res_ps is not None:
        from dataclasses import replace
        sc = replace(scanner, tof_res_ps=tof_res_ps)
    return tof_events(n, _ImagePhantom(), seed=seed, scanner=sc, device=device, batch=batch,
                      gen=gen)


def lumpy_image(grid, center, size, seed, n_blobs=50, cyl_radius=150.0, device="cpu"):
    """Sum of n_blobs Gaussians (sigma uniform in [5, 20] mm, centres uniform in
    the cylinder of radius cyl_radius and the image's axial extent)."""
    import torch
    rng = np.random.default_rng(seed)
    axes = [center[k] - size[k] / 2 + (np.arange(grid[k]) + 0.5) * size[k] / grid[k]
            for k in range(3)]
    X, Y, Z = [torch.tensor(a, dtype=torch.float64, device=device) for a in axes]
    img = torch.zeros(grid, dtype=torch.float64, device=device)
    for _ in range(n_blobs):
        r = cyl_radius * np.sqrt(rng.uniform())
        ph = rng.uniform(0, 2 * np.pi)
        c = (r * np.cos(ph), r * np.sin(ph), rng.uniform(-size[2] / 2, size[2] / 2))
        s = rng.uniform(5, 20)
        a = rng.uniform(0.5, 2.0)
        gx = torch.exp(-(X - c[0]) ** 2 / (2 * s * s))
        gy = torch.exp(-(Y - c[1]) ** 2 / (2 * s * s))
        gz = torch.exp(-(Z - c[2]) ** 2 / (2 * s * s))
        img += a * gx[:, None, None] * gy[None, :, None] * gz[None, None, :]
    return img.to(torch.float32)


def torso_image(grid, center, size, seed=5, device="cpu"):
    """Torso-like activity and mu-map (nx, ny, nz): elliptic body 350 x 250 mm
    (activity 1, mu 0.0096/mm), two lungs (activity 0.3), liver (1.5) and
    10 spherical lesions of 5-15 mm radius at 3-6x the background."""
    import torch
    rng = np.random.default_rng(seed)
    axes = [center[k] - size[k] / 2 + (np.arange(grid[k]) + 0.5) * size[k] / grid[k]
            for k in range(3)]
    X, Y, Z = torch.meshgrid(*[torch.tensor(a, dtype=torch.float64, device=device)
                               for a in axes], indexing="ij")
    body = (X / 175.0) ** 2 + (Y / 125.0) ** 2 <= 1.0
    act = body.to(torch.float64)
    for cx in (-80.0, 80.0):
        lung = ((X - cx) / 55.0) ** 2 + ((Y - 10.0) / 75.0) ** 2 + (Z / 120.0) ** 2 <= 1.0
        act = torch.where(lung, torch.full_like(act, 0.3), act)
    liver = ((X + 60.0) / 80.0) ** 2 + ((Y + 20.0) / 60.0) ** 2 + ((Z + 90.0) / 60.0) ** 2 <= 1
    act = torch.where(liver & body, torch.full_like(act, 1.5), act)
    for _ in range(10):
        r = rng.uniform(5, 15)
        while True:
            c = (rng.uniform(-150, 150), rng.uniform(-100, 100), rng.uniform(-100, 100))
            if (c[0] / 175) ** 2 + (c[1] / 125) ** 2 < 0.8:
                break
        les = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 <= r * r
        act = torch.where(les, torch.full_like(act, rng.uniform(3, 6)), act)
    mu = torch.where(body, torch.full_like(act, 0.0096), torch.zeros_like(act))
    return act.to(torch.float32), mu.to(torch.float32)
