"""TEST INFRASTRUCTURE ONLY — numpy / scipy restatement of the PSF blur of
PSFReconStep (src/python/srf/model/recon_step.py:56-115), in FP64.  Images
are (nx, ny, nz) arrays in C order; psf_xy is an (nx*ny) x (nx*ny) scipy
sparse matrix, psf_z an nz x nz array."""
import numpy as np


def psf_forward(image, psf_xy, psf_z):
    """recon_step.py:77-90: reshape to [nx*ny, nz], @ psf_z^T, then the sparse
    psf_xy @ (.), reshape back."""
    grid = image.shape
    image_vectorized = np.asarray(image, np.float64).reshape(grid[0] * grid[1], grid[2])
    image_mul_psf_z = image_vectorized @ np.asarray(psf_z, np.float64).T
    image_psf_processed = psf_xy @ image_mul_psf_z
    return np.asarray(image_psf_processed).reshape(grid)


def psf_backward(back_proj, psf_xy, psf_z):
    """recon_step.py:103-110: reshape, @ psf_z, then psf_xy^T @ (.)."""
    grid = back_proj.shape
    back_proj_vectorized = np.asarray(back_proj, np.float64).reshape(grid[0] * grid[1], grid[2])
    back_proj_mul_psf_z = back_proj_vectorized @ np.asarray(psf_z, np.float64)
    back_proj_psf_processed = psf_xy.T @ back_proj_mul_psf_z
    return np.asarray(back_proj_psf_processed).reshape(grid)
