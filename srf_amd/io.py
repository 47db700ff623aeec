"""Files of the reference's reconstruction workflow, read straight into device
memory.

load_lors            list-mode rows (N, 7) = x1, y1, z1, x2, y2, z2, tof from
                     .npy (SplitWorkerLoader.load, data/loader.py:97-99) or
                     from the HDF5 group ``listmode_data`` (columns fst, snd,
                     tof; io/listmode.py:18-25, loader.py:100-104).  .npy
                     files are memory-mapped and streamed to the device in
                     chunks through two pinned staging buffers, so a file
                     larger than host memory (300M events: 8.4 GB as float32,
                     16.8 GB as float64) never needs a host copy.
save_lors            (N, 7) rows to .npy.
load_image / save_image
                     (nx, ny, nz) float32 images as .npy: the efficiency map
                     file (loader.py:112-114) and the per-iteration outputs
                     (local_reconstruction.py:51, ``{prefix}_{i}.npy``).
split_worker_data    SplitWorkerLoader.load (loader.py:96-119): rows ->
                     recon_process on the device -> ListModeDataSplit (values
                     1) + the efficiency-map Image.
complete_worker_data CompleteWorkerLoader.load (loader.py:122-145): rows ->
                     ListModeData(rows, rows[:, 6]) for the Siddon model +
                     the efficiency-map Image.

HDF5 needs h5py, which this environment does not provide: the HDF5 branch
raises ImportError here and is exercised only where h5py is installed.
"""
import numpy as np
import torch

from .data import Image, ListModeData, split_from_groups
from .preprocess import recon_process

H5_GROUP = "listmode_data"


def _stream_npy(path, device, dtype, chunk_rows):
    mm = np.load(path, mmap_mode="r", allow_pickle=False)
    if mm.ndim != 2 or mm.shape[1] != 7:
        raise ValueError(f"{path}: expected (N, 7) list-mode rows, got {mm.shape}")
    n = mm.shape[0]
    out = torch.empty((n, 7), dtype=dtype, device=device)
    if n == 0:
        return out
    host_dtype = torch.from_numpy(np.empty(0, dtype=mm.dtype)).dtype
    rows = max(1, min(int(chunk_rows), n))
    if torch.device(device).type != "cuda":
        for a in range(0, n, rows):
            b = min(n, a + rows)
            out[a:b] = torch.from_numpy(np.array(mm[a:b])).to(dtype)
        return out
    stage = [torch.empty((rows, 7), dtype=host_dtype, pin_memory=True) for _ in range(2)]
    done = [None, None]
    stream = torch.cuda.current_stream(device)
    for k, a in enumerate(range(0, n, rows)):
        b = min(n, a + rows)
        buf = stage[k % 2]
        if done[k % 2] is not None:
            done[k % 2].synchronize()  # the copy that last read this buffer has ended
        buf.numpy()[:b - a] = mm[a:b]  # file -> pinned buffer, one host copy
        out[a:b].copy_(buf[:b - a], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(stream)
        done[k % 2] = ev
    torch.cuda.current_stream(device).synchronize()
    return out


def _load_h5(path):
    try:
        import h5py
    except ImportError as err:
        raise ImportError(f"{path}: reading HDF5 list-mode files needs h5py") from err
    with h5py.File(path, "r") as fin:
        group = fin[H5_GROUP]
        return {k: np.array(group[k]) for k in group.keys()}


def load_lors(path, device="cpu", dtype=torch.float64, chunk_rows=1 << 22):
    """(N, 7) list-mode rows of ``path`` (.npy or .h5) as a ``dtype`` tensor on
    ``device``."""
    path = str(path)
    if path.endswith(".npy"):
        return _stream_npy(path, device, dtype, chunk_rows)
    if path.endswith(".h5"):
        d = _load_h5(path)
        rows = np.hstack([d["fst"], d["snd"], d["tof"].reshape(d["tof"].size, 1)])
        return torch.from_numpy(rows).to(device=device, dtype=dtype)
    raise ValueError(f"{path}: list-mode files are .npy or .h5")


def save_lors(path, lors):
    lors = lors.detach().cpu().numpy() if torch.is_tensor(lors) else np.asarray(lors)
    if lors.ndim != 2 or lors.shape[1] != 7:
        raise ValueError(f"expected (N, 7) list-mode rows, got {lors.shape}")
    np.save(path, lors)


def load_image(path, center, size, device="cpu"):
    """Image of a (nx, ny, nz) .npy file as float32 on ``device``
    (loader.py:112-114 reads the efficiency map this way)."""
    a = np.load(str(path), allow_pickle=False)
    if a.ndim != 3:
        raise ValueError(f"{path}: expected an (nx, ny, nz) image, got {a.shape}")
    return Image(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device),
                 center, size)


def save_image(path, image):
    """np.save of the (nx, ny, nz) data (local_reconstruction.py:51)."""
    data = image.data if isinstance(image, Image) else image
    np.save(str(path), data.detach().cpu().numpy() if torch.is_tensor(data) else data)


def split_worker_data(lors_path, emap_path, center, size, tof_res, device):
    """SplitWorkerLoader.load: {'projection_data': ListModeDataSplit,
    'efficiency_map': Image}, the events preprocessed on ``device``."""
    rows = load_lors(lors_path, device)
    groups = recon_process(rows, tof_res)
    return {"projection_data": split_from_groups(groups),
            "efficiency_map": load_image(emap_path, center, size, device)}


def complete_worker_data(lors_path, emap_path, center, size, device):
    """CompleteWorkerLoader.load: {'projection_data': ListModeData(rows,
    rows[:, 6]), 'efficiency_map': Image} (float32 rows: x1..z2, tof)."""
    rows = load_lors(lors_path, device, dtype=torch.float32)
    return {"projection_data": ListModeData(rows, rows[:, 6].clone()),
            "efficiency_map": load_image(emap_path, center, size, device)}
