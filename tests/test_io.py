"""List-mode and image files (srf_amd.io): .npy round trips, the chunked
streaming reader, the worker loaders against the oracle preprocessing, and
the HDF5 branch's error without h5py."""
import importlib.util

import numpy as np
import pytest
import torch

from srf_amd import io


def _rows(n, seed, dtype):
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * np.pi, (n, 2))
    z = rng.uniform(-50, 50, (n, 2))
    rows = np.stack([400 * np.cos(ang[:, 0]), 400 * np.sin(ang[:, 0]), z[:, 0],
                     400 * np.cos(ang[:, 1]), 400 * np.sin(ang[:, 1]), z[:, 1],
                     rng.normal(0, 30, n)], axis=1)
    return rows.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("chunk_rows", [7, 1000, 1 << 22])
def test_lors_round_trip_chunked(tmp_path, dtype, chunk_rows):
    rows = _rows(1003, 0, dtype)
    path = tmp_path / "lors.npy"
    io.save_lors(path, rows)
    got = io.load_lors(path, "cpu", dtype=torch.float64, chunk_rows=chunk_rows)
    np.testing.assert_array_equal(got.numpy(), rows.astype(np.float64))
    empty = tmp_path / "empty.npy"
    io.save_lors(empty, np.zeros((0, 7), dtype))
    assert io.load_lors(empty, "cpu").shape == (0, 7)


def test_bad_files(tmp_path):
    np.save(tmp_path / "bad.npy", np.zeros((4, 6)))
    with pytest.raises(ValueError):
        io.load_lors(tmp_path / "bad.npy")
    with pytest.raises(ValueError):
        io.load_lors(tmp_path / "lors.txt")
    with pytest.raises(ValueError):
        io.save_lors(tmp_path / "x.npy", np.zeros((3, 6)))


@pytest.mark.skipif(importlib.util.find_spec("h5py") is not None, reason="h5py present")
def test_h5_needs_h5py(tmp_path):
    with pytest.raises(ImportError):
        io.load_lors(tmp_path / "lors.h5")


def test_split_worker_data_matches_oracle(tmp_path):
    from oracle import preprocess as OP
    rows = _rows(2000, 1, np.float64)
    np.save(tmp_path / "lors.npy", rows)
    emap = np.random.default_rng(2).uniform(0.5, 2, (8, 9, 5)).astype(np.float32)
    np.save(tmp_path / "emap.npy", emap)
    d = io.split_worker_data(tmp_path / "lors.npy", tmp_path / "emap.npy", (0, 0, 0),
                             (400, 400, 110), 530.0, "cpu")
    ref = OP.recon_process(rows, 530.0)
    for a in "xyz":
        np.testing.assert_array_equal(d["projection_data"][a].lors.numpy(),
                                      ref[a].astype(np.float32))
        assert torch.all(d["projection_data"][a].values == 1)
    e = d["efficiency_map"]
    np.testing.assert_array_equal(e.data.numpy(), emap)
    assert e.center == [0, 0, 0] and e.size == [400, 400, 110] and e.grid == [8, 9, 5]


def test_complete_worker_data_and_images(tmp_path):
    rows = _rows(50, 3, np.float32)
    np.save(tmp_path / "lors.npy", rows)
    img = np.arange(60, dtype=np.float32).reshape(3, 4, 5)
    io.save_image(tmp_path / "recon_0.npy", torch.from_numpy(img))
    d = io.complete_worker_data(tmp_path / "lors.npy", tmp_path / "recon_0.npy", (0, 0, 0),
                                (30, 40, 50), "cpu")
    np.testing.assert_array_equal(d["projection_data"].lors.numpy(), rows)
    np.testing.assert_array_equal(d["projection_data"].values.numpy(), rows[:, 6])
    np.testing.assert_array_equal(d["efficiency_map"].data.numpy(), img)
    with pytest.raises(ValueError):
        np.save(tmp_path / "flat.npy", np.zeros(5))
        io.load_image(tmp_path / "flat.npy", (0, 0, 0), (1, 1, 1))


@pytest.mark.gpu
def test_stream_to_device(cuda, tmp_path):
    rows = _rows(100_003, 4, np.float32)
    np.save(tmp_path / "lors.npy", rows)
    for chunk in (999, 1 << 22):
        got = io.load_lors(tmp_path / "lors.npy", cuda, dtype=torch.float64, chunk_rows=chunk)
        assert got.device.type == "cuda"
        np.testing.assert_array_equal(got.cpu().numpy(), rows.astype(np.float64))
