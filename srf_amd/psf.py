"""PSF-modelled reconstruction step (PSFReconStep, model/recon_step.py:56-115).

PSF(psf_xy, psf_z) holds the transaxial blur (a sparse nxy x nxy matrix, nxy =
nx * ny, over the image's voxel columns in C order) and the axial blur (a
dense nz x nz matrix) on the device:

  forward(image)   reshape(x, [nxy, nz]) @ psf_z^T, then psf_xy @ (.)
                   (recon_step.py:79-90): the image the projector sees.
  backward(image)  reshape(b, [nxy, nz]) @ psf_z, then psf_xy^T @ (.)
                   (recon_step.py:103-110): applied to the backprojection.

Both run as one HIP kernel (srfx_psf_apply).  load_psf reads the files of
WorkerLoader._load_psf (data/loader.py:75-86): psf_xy from a MATLAB file's
'matrix' entry (scipy.io.loadmat), psf_z from .npy.  gaussian_psf builds
separable Gaussian kernels for tests and benchmarks (the reference's
PSFMaker fits its kernels from measured point sources, which is outside this
package).  PSFReconStep / OSEM(psf=...) in recon.py use them.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .data import Image
from .ops import _check_dev, _ptr, _stream


def _bind():
    lib = _lib.load()
    if not getattr(lib, "_psf_bound", False):
        lib.srfx_psf_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
        lib.srfx_psf_apply.restype = ctypes.c_int
        lib._psf_bound = True
    return lib


def _band(m):
    """(below, above): m[z', z] == 0 unless z - below <= z' <= z + above."""
    zp, z = np.nonzero(m)
    if zp.size == 0:
        return 0, 0
    return int(max(0, np.max(z - zp))), int(max(0, np.max(zp - z)))


def _csr(mat, n):
    """(row_ptr, col, val) host arrays of an n x n scipy sparse or dense matrix."""
    from scipy import sparse
    m = sparse.csr_matrix(mat, dtype=np.float64)
    if m.shape != (n, n):
        raise ValueError(f"psf_xy must be {n} x {n}, got {m.shape}")
    m.sum_duplicates()
    m.sort_indices()
    if m.nnz > np.iinfo(np.int32).max:
        raise ValueError("psf_xy has more than 2^31 - 1 entries")
    return (m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data.astype(np.float32))


class PSF:
    def __init__(self, psf_xy, psf_z, grid, device):
        """psf_xy: (nx*ny) x (nx*ny) scipy sparse or dense; psf_z: nz x nz;
        grid: (nx, ny, nz)."""
        from scipy import sparse
        self.grid = tuple(int(g) for g in grid)
        nxy, nz = self.grid[0] * self.grid[1], self.grid[2]
        z = np.asarray(psf_z, dtype=np.float64)
        if z.shape != (nz, nz):
            raise ValueError(f"psf_z must be {nz} x {nz}, got {z.shape}")
        self.device = torch.device(device)
        xy = sparse.csr_matrix(psf_xy, dtype=np.float64)
        self._fwd = [torch.from_numpy(a).to(self.device) for a in _csr(xy, nxy)]
        self._bwd = [torch.from_numpy(a).to(self.device) for a in _csr(xy.T, nxy)]
        mf, mb = np.ascontiguousarray(z.T, dtype=np.float32), np.ascontiguousarray(z, np.float32)
        self._band_f, self._band_b = _band(mf), _band(mb)
        self._mf = torch.from_numpy(mf).to(self.device)
        self._mb = torch.from_numpy(mb).to(self.device)

    def _apply(self, csr, m, band, x, out):
        _check_dev(x, "image")
        if tuple(x.shape) != self.grid or x.dtype != torch.float32 or not x.is_contiguous():
            raise ValueError(f"image must be a contiguous float32 {self.grid} tensor")
        if out is None:
            out = torch.empty_like(x)
        lib = _bind()
        row_ptr, col, val = csr
        st = lib.srfx_psf_apply(_ptr(row_ptr), _ptr(col), _ptr(val), self.grid[0] * self.grid[1],
                                _ptr(m), self.grid[2], band[0], band[1], _ptr(x), _ptr(out),
                                _stream())
        if st != _lib.SRFX_OK:
            raise _lib.SrfxError("srfx_psf_apply", st, lib.srfx_strerror(st).decode())
        return out

    def forward(self, image, out=None):
        """psf_xy (x psf_z^T), as an Image (recon_step.py:79-91)."""
        return Image(self._apply(self._fwd, self._mf, self._band_f, image.data, out),
                     image.center, image.size)

    def backward(self, image, out=None):
        """psf_xy^T (b psf_z), as an Image (recon_step.py:103-113)."""
        return Image(self._apply(self._bwd, self._mb, self._band_b, image.data, out),
                     image.center, image.size)


def load_psf(xy_path, z_path, grid, device):
    """WorkerLoader._load_psf (loader.py:75-86): psf_xy = loadmat(xy)['matrix'],
    psf_z = np.load(z)."""
    import scipy.io as scio
    from scipy import sparse
    xy = sparse.coo_matrix(scio.loadmat(str(xy_path))["matrix"])
    return PSF(xy, np.load(str(z_path), allow_pickle=False), grid, device)


def gaussian_psf(grid, size, fwhm_xy, fwhm_z, device, cutoff=3.0):
    """Separable Gaussian blur: transaxial sigma fwhm_xy / 2.3548 mm over voxel
    centres within cutoff sigmas (rows normalised to 1), axial likewise."""
    from scipy import sparse
    nx, ny, nz = (int(g) for g in grid)
    vx, vy, vz = (size[k] / grid[k] for k in range(3))
    sxy, sz = fwhm_xy / 2.35482005, fwhm_z / 2.35482005
    rx, ry = int(np.ceil(cutoff * sxy / vx)), int(np.ceil(cutoff * sxy / vy))
    offs = [(dx, dy, np.exp(-((dx * vx) ** 2 + (dy * vy) ** 2) / (2 * sxy * sxy)))
            for dx in range(-rx, rx + 1) for dy in range(-ry, ry + 1)
            if (dx * vx) ** 2 + (dy * vy) ** 2 <= (cutoff * sxy) ** 2]
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    rows, cols, vals = [], [], []
    for dx, dy, w in offs:
        jx, jy = ix + dx, iy + dy
        ok = (jx >= 0) & (jx < nx) & (jy >= 0) & (jy < ny)
        rows.append((ix * ny + iy)[ok])
        cols.append((jx * ny + jy)[ok])
        vals.append(np.full(ok.sum(), w))
    xy = sparse.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                           shape=(nx * ny, nx * ny))
    xy = sparse.diags(1.0 / np.asarray(xy.sum(axis=1)).ravel()) @ xy
    k = np.arange(nz)
    z = np.exp(-((k[:, None] - k[None, :]) * vz) ** 2 / (2 * sz * sz))
    z[np.abs(k[:, None] - k[None, :]) * vz > cutoff * sz] = 0.0
    z /= z.sum(axis=1, keepdims=True)
    return PSF(xy, z, grid, device)
