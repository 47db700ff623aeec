"""PSF blur (PSFReconStep, recon_step.py:56-115): the oracle restatement pinned
by the Kronecker form vec(S X M) = (S kron M^T) vec(X) of a row-major image
matrix and by adjointness; the host-side kernel builders and file loader."""
import numpy as np
import pytest
import scipy.io as scio
from scipy import sparse

from oracle import psf as OPSF


def _mats(grid, seed, density=0.3):
    rng = np.random.default_rng(seed)
    nxy, nz = grid[0] * grid[1], grid[2]
    xy = sparse.random(nxy, nxy, density=density, random_state=seed, format="csr")
    z = rng.uniform(-1, 1, (nz, nz))  # not symmetric: a transposed operand shows
    return xy, z


def test_oracle_forward_is_kronecker():
    grid = (3, 4, 5)
    xy, z = _mats(grid, 0)
    x = np.random.default_rng(1).normal(size=grid)
    got = OPSF.psf_forward(x, xy, z).reshape(-1)
    # forward = psf_xy X psf_z^T: vec_row(A X B) = (A kron B^T) vec_row(X), B = psf_z^T
    want = np.kron(xy.toarray(), z) @ x.reshape(-1)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_oracle_backward_is_adjoint():
    grid = (4, 3, 6)
    xy, z = _mats(grid, 2)
    rng = np.random.default_rng(3)
    x, y = rng.normal(size=grid), rng.normal(size=grid)
    lhs = np.sum(OPSF.psf_forward(x, xy, z) * y)
    rhs = np.sum(x * OPSF.psf_backward(y, xy, z))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    want = np.kron(xy.toarray().T, z.T) @ y.reshape(-1)
    np.testing.assert_allclose(OPSF.psf_backward(y, xy, z).reshape(-1), want, rtol=1e-12,
                               atol=1e-12)


def test_gaussian_psf_rows_normalised():
    from srf_amd.psf import gaussian_psf
    grid, size = (9, 7, 6), (90.0, 70.0, 66.0)
    p = gaussian_psf(grid, size, 25.0, 20.0, "cpu")
    row_ptr, col, val = (t.numpy() for t in p._fwd)
    sums = np.add.reduceat(val.astype(np.float64), row_ptr[:-1])
    np.testing.assert_allclose(sums, 1.0, rtol=1e-6)
    assert col.min() >= 0 and col.max() < 9 * 7
    np.testing.assert_allclose(p._mb.numpy().sum(axis=1), 1.0, rtol=1e-6)
    np.testing.assert_array_equal(p._mf.numpy(), p._mb.numpy().T)
    # interior voxel column: the full disc of offsets
    r = 4 * 7 + 3
    assert row_ptr[r + 1] - row_ptr[r] == np.max(np.diff(row_ptr))


def test_load_psf_files(tmp_path):
    from srf_amd.psf import load_psf
    grid = (3, 4, 5)
    xy, z = _mats(grid, 4)
    scio.savemat(tmp_path / "psf_xy.mat", {"matrix": xy})
    np.save(tmp_path / "psf_z.npy", z)
    p = load_psf(tmp_path / "psf_xy.mat", tmp_path / "psf_z.npy", grid, "cpu")
    row_ptr, col, val = (t.numpy() for t in p._fwd)
    dense = sparse.csr_matrix((val, col, row_ptr), shape=(12, 12)).toarray()
    np.testing.assert_allclose(dense, xy.toarray().astype(np.float32))
    row_ptr, col, val = (t.numpy() for t in p._bwd)
    dense_t = sparse.csr_matrix((val, col, row_ptr), shape=(12, 12)).toarray()
    np.testing.assert_allclose(dense_t, xy.toarray().T.astype(np.float32))
    np.testing.assert_allclose(p._mf.numpy(), z.T.astype(np.float32))
    with pytest.raises(ValueError):
        load_psf(tmp_path / "psf_xy.mat", tmp_path / "psf_z.npy", (3, 4, 6), "cpu")


def test_band_of_z_matrix():
    from srf_amd.psf import _band
    m = np.zeros((6, 6))
    m[2, 4] = 1.0  # z' = 2 feeds z = 4: below 2
    m[5, 3] = 1.0  # z' = 5 feeds z = 3: above 2
    m[1, 1] = 1.0
    assert _band(m) == (2, 2)
    assert _band(np.eye(5)) == (0, 0)
    assert _band(np.ones((4, 4))) == (3, 3)
    assert _band(np.triu(np.ones((4, 4)))) == (3, 0)
