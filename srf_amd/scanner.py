"""Cylindrical PET scanner geometry: crystal centres and crystal-pair LORs.

Follows the reference's CylindricalPET / RingBlock construction
(src/python/srf/scanner/pet/pet.py:97-177, block.py:77-108) for blocks of a
single crystal: ring r sits at z = -(pitch + gap)(R-1)/2 + r (pitch + gap) and
block b of a ring at angle 2 pi b / B about the z axis, radius
(inner + outer)/2.
"""
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class CylindricalScanner:
    inner_radius: float = 390.0
    outer_radius: float = 410.0
    nb_rings: int = 32
    nb_crystals: int = 384          # blocks per ring, one crystal each
    ring_pitch: float = 3.4         # block size along z
    gap: float = 0.0
    tof_res_ps: float = 530.0
    tof_bin_ps: float = 40.0

    @property
    def radius(self):
        return (self.inner_radius + self.outer_radius) / 2

    @property
    def axial_length(self):
        return self.nb_rings * (self.ring_pitch + self.gap) - self.gap

    def ring_z(self):
        step = self.ring_pitch + self.gap
        bottom = -step * (self.nb_rings - 1) / 2
        return bottom + step * np.arange(self.nb_rings)

    def crystal_angles(self):
        return np.arange(self.nb_crystals) * 2 * np.pi / self.nb_crystals

    def crystal_xyz(self, ring, crystal):
        """Centres of crystals (ring, crystal) (broadcasting index arrays)."""
        ang = self.crystal_angles()[np.asarray(crystal)]
        z = self.ring_z()[np.asarray(ring)]
        r = self.radius
        return np.stack(np.broadcast_arrays(r * np.cos(ang), r * np.sin(ang), z), axis=-1)

    def tof_info(self):
        """(tof_bin [mm], tof_sigma2 [mm^2]) as SRFApp._get_tof_info derives
        them (srfapp_new.py:244-251)."""
        gaussian_factor = 2.35482005
        limit = self.tof_res_ps * 0.15 / gaussian_factor * 3
        return self.tof_bin_ps * 0.15, (limit ** 2) / 9

    def all_pairs(self, chunk=None):
        """Every unordered crystal pair once, as (N, 6) float32 end points
        (the efficiency-map LOR set).  With chunk=(start, stop) only pair
        indices in that range of the fixed enumeration are built.

        Enumeration: crystal ids k = ring * nb_crystals + crystal, pairs
        (a, b) with a < b in lexicographic order."""
        nc = self.nb_rings * self.nb_crystals
        total = nc * (nc - 1) // 2
        start, stop = (0, total) if chunk is None else chunk
        idx = np.arange(start, stop, dtype=np.int64)
        # invert the lexicographic pair index: a = number of full rows before idx
        # row a holds (nc - 1 - a) pairs; offset(a) = a*nc - a*(a+1)/2
        a = np.floor((2 * nc - 1 - np.sqrt((2 * nc - 1) ** 2 - 8.0 * idx)) / 2).astype(np.int64)
        off = a * nc - a * (a + 1) // 2
        fix = idx < off
        a[fix] -= 1
        off = a * nc - a * (a + 1) // 2
        fix = idx >= off + (nc - 1 - a)
        a[fix] += 1
        off = a * nc - a * (a + 1) // 2
        b = idx - off + a + 1
        pa = self.crystal_xyz(a // self.nb_crystals, a % self.nb_crystals)
        pb = self.crystal_xyz(b // self.nb_crystals, b % self.nb_crystals)
        return np.hstack([pa, pb]).astype(np.float32)

    def crystal_table(self):
        """(nb_rings * nb_crystals, 3) float32 centres, indexed by crystal id
        k = ring * nb_crystals + crystal (the values all_pairs uses)."""
        k = np.arange(self.nb_rings * self.nb_crystals)
        return self.crystal_xyz(k // self.nb_crystals, k % self.nb_crystals).astype(np.float32)

    def all_pairs_device(self, chunk, device):
        """all_pairs(chunk) built on ``device`` (torch, float32 [N, 6]); the
        same pairs in the same order and the same float32 values."""
        import torch
        nc = self.nb_rings * self.nb_crystals
        start, stop = chunk
        idx = torch.arange(start, stop, dtype=torch.int64, device=device)
        q = (2 * nc - 1) ** 2 - 8.0 * idx.double()
        a = torch.floor((2 * nc - 1 - torch.sqrt(q)) / 2).long()

        def off(a):
            return a * nc - a * (a + 1) // 2
        a = torch.where(idx < off(a), a - 1, a)
        a = torch.where(idx >= off(a) + (nc - 1 - a), a + 1, a)
        b = idx - off(a) + a + 1
        table = torch.from_numpy(self.crystal_table()).to(device)
        return torch.cat([table[a], table[b]], dim=1)

    def ring_pair_lors(self, ring1, ring2, device="cpu"):
        """PETScanner.make_ring_pairs_lors (pet.py:151-177) for single-crystal
        blocks: the crystals i1 of ring1 and i2 of ring2 with i1 < i2 when the
        rings are the same, i1 != i2 otherwise, i1 outer and i2 inner, as
        (N, 6) float32 end points [ring1 crystal, ring2 crystal] (torch, on
        ``device``)."""
        import torch
        b = self.nb_crystals
        i1 = torch.arange(b, device=device).repeat_interleave(b)
        i2 = torch.arange(b, device=device).repeat(b)
        keep = i1 < i2 if ring1 == ring2 else i1 != i2
        table = torch.from_numpy(self.crystal_table()).to(device)
        return torch.cat([table[ring1 * b + i1[keep]], table[ring2 * b + i2[keep]]], dim=1)

    def nb_pairs(self):
        nc = self.nb_rings * self.nb_crystals
        return nc * (nc - 1) // 2
