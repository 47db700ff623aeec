"""List-mode MLEM / OSEM driver.

Follows the reference's semantics:
  ReconStep.kernel + mlem_update   x <- x * E * sum_groups A^T(1 / A x)
                                    (src/python/srf/model/recon_step.py:33-49)
  subsets                          contiguous chunks of each group's rows,
                                    subset s of K = rows [s*N/K, (s+1)*N/K)
                                    (graph/reconstruction/worker.py:79-85)
  subset counter                   (s + 1) % K (master.py:97-111)
  iterations                       LocalReconstructionGraph.run
                                    (local_reconstruction.py:44-51)
E is the stored inverse efficiency map and is applied unscaled to every
subset, as the reference does.

PSFReconStep (recon_step.py:56-115): with a psf.PSF, the projector sees the
blurred image PSF(x) and the backprojection is blurred back with PSF^T before
the update, x <- x * E * PSF^T(sum_groups A^T(1 / A PSF(x))).  ReconStep and
OSEM take it as ``psf=``.
"""
import torch

from . import ops
from .data import Image, ListModeData, ListModeDataSplit


def subset_bounds(n, k, nb_subsets):
    """Rows of subset k: [k*n//K, (k+1)*n//K).  For n divisible by K this is
    the reference's contiguous split (test_worker.py:105-121 pins N=30, K=3)."""
    return (k * n) // nb_subsets, ((k + 1) * n) // nb_subsets


def _subset_part(lm, k, nb_subsets):
    lo, hi = subset_bounds(len(lm), k, nb_subsets)
    part = lm.replace_values(lm.values[lo:hi] if lm.values is not None else None)
    part.soa = lm.soa[:, lo:hi]
    return part


def subset_view(data, k, nb_subsets):
    """Column views (no copy) of each group's SoA LORs for subset k (a
    ListModeDataSplit of the TOR model), or of the LORs of a single
    ListModeData (the Siddon model's complete LORs)."""
    if isinstance(data, ListModeData):
        return _subset_part(data, k, nb_subsets)
    parts = []
    for a in ("x", "y", "z"):
        lm = data[a]
        lo, hi = subset_bounds(len(lm), k, nb_subsets)
        v = lm.values[lo:hi] if lm.values is not None else None
        part = lm.replace_values(v)
        part.soa = lm.soa[:, lo:hi]
        parts.append(part)
    return ListModeDataSplit(*parts)


def mlem_update(image_prev, image_succ, efficiency_map):
    """recon_step.py:48-49: x * E * bp, in place on image_prev.data."""
    ops.mlem_update(image_prev.data, efficiency_map.data, image_succ.data)
    return image_prev


class ReconStep:
    """One reconstruction step with the reference's structure
    (recon_step.py:17-45): projection, ratio backprojection, update."""

    def __init__(self, model, fused=True, psf=None):
        self.model = model
        self.fused = fused
        self.psf = psf
        self._bp = None
        self._blur = None

    def __call__(self, image, data, efficiency_map, plans=None):
        if self._bp is None or self._bp.shape != image.data.shape:
            self._bp = torch.empty_like(image.data)
            self._blur = torch.empty_like(image.data) if self.psf is not None else None
        self._bp.zero_()
        seen = image if self.psf is None else self.psf.forward(image, out=self._blur)
        if self.fused:
            if plans is None:
                self.model.osem_subset(seen, data, out=self._bp)
            else:
                self.model.osem_subset(seen, data, out=self._bp, plans=plans)
        else:
            proj = self.model.projection(seen, data)
            if plans is None:
                self.model.backprojection(proj, seen, out=self._bp)
            else:
                self.model.backprojection(proj, seen, out=self._bp, plans=plans)
        bp = Image(self._bp, image.center, image.size)
        if self.psf is not None:
            bp = self.psf.backward(bp, out=self._blur)  # the blurred image is no longer needed
        return mlem_update(image, bp, efficiency_map)


def PSFReconStep(model, psf, fused=True):
    """recon_step.py:56-115: ReconStep with the PSF blur around the projector."""
    return ReconStep(model, fused=fused, psf=psf)


class OSEM:
    """iterations x subsets of ReconStep over list-mode data split in groups.

    tiled=True (models with make_plans, the TOR model): the events of every
    subset are binned into image tiles once (``plan``, on the first subset
    run or explicitly) and every later step of that subset reuses the plan."""

    def __init__(self, model, data, efficiency_map, nb_subsets=1, fused=True, tiled=False,
                 psf=None):
        self.step = ReconStep(model, fused=fused, psf=psf)
        self.model = model
        self.data = data
        self.emap = efficiency_map
        self.nb_subsets = int(nb_subsets)
        self.subsets = [subset_view(data, k, self.nb_subsets) for k in range(self.nb_subsets)]
        self.subset = 0
        self.tiled = bool(tiled) and hasattr(model, "make_plans")
        self.plans = None

    def plan(self, image):
        """Build the tile plans of every subset for images laid out like image."""
        self.close()
        if hasattr(self.model, "make_subset_plans"):
            # every subset of a group in one build (two synchronisations)
            self.plans = self.model.make_subset_plans(image, self.data, self.nb_subsets)
        else:
            self.plans = [self.model.make_plans(image, sub) for sub in self.subsets]
        return self.plans

    def close(self):
        for ps in self.plans or []:
            for p in ps.values():
                if p is not None:
                    p.close()
        self.plans = None

    def run_subset(self, image):
        plans = None
        if self.tiled:
            if self.plans is None:
                self.plan(image)
            plans = self.plans[self.subset]
        self.step(image, self.subsets[self.subset], self.emap, plans=plans)
        self.subset = (self.subset + 1) % self.nb_subsets
        return image

    def iterate(self, image, nb_iterations=1, callback=None):
        for it in range(nb_iterations):
            for _ in range(self.nb_subsets):
                self.run_subset(image)
            if callback is not None:
                callback(it, image)
        return image
