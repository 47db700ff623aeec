"""Data carriers mirroring src/python/srf/data/{image,listmode}.py.

Image            data (nx, ny, nz) tensor + center + size (full extent, mm);
                 grid = data.shape (image.py:4-28).
ListModeData     LORs + per-LOR values (listmode.py:52-66).  LORs are held as
                 a [ncols, N] structure-of-arrays device tensor; ``lors``
                 returns the reference's [N, ncols] view of it.
ListModeDataSplit / ListModeDataSplitWithoutTOF
                 one ListModeData per axis group x, y, z (listmode.py:118-138).
"""
import torch


class Image:
    def __init__(self, data, center, size):
        self.data = data
        self.center = [float(c) for c in center]
        self.size = [float(s) for s in size]

    @property
    def grid(self):
        return list(self.data.shape)

    def fmap(self, f):
        return Image(f(self.data), self.center, self.size)


class ListModeData:
    def __init__(self, lors, values=None, soa=False):
        """lors: [N, ncols] rows (reference layout) or, with soa=True, the
        [ncols, N] structure-of-arrays tensor used by the ops."""
        self.soa = lors.contiguous() if soa else lors.t().contiguous()
        self.values = values

    @property
    def lors(self):
        return self.soa.t()

    def __len__(self):
        return self.soa.shape[1]

    def replace_values(self, values):
        out = ListModeData.__new__(ListModeData)
        out.soa = self.soa
        out.values = values
        return out


class ListModeDataWithoutTOF(ListModeData):
    pass


class _Split:
    def __init__(self, x, y, z):
        self.x, self.y, self.z = x, y, z

    def __getitem__(self, item):
        if item in ("x", "y", "z"):
            return getattr(self, item)
        raise KeyError(item)

    def items(self):
        return (("x", self.x), ("y", self.y), ("z", self.z))


class ListModeDataSplit(_Split):
    pass


class ListModeDataSplitWithoutTOF(_Split):
    pass


def split_from_groups(groups, values=None, without_tof=False):
    """Wrap preprocess output {'x','y','z'}: [ncols, N_a] SoA tensors."""
    cls = ListModeDataWithoutTOF if without_tof else ListModeData
    parts = []
    for a in ("x", "y", "z"):
        g = groups[a]
        v = values[a] if values is not None else torch.ones(g.shape[1], dtype=torch.float32,
                                                            device=g.device)
        parts.append(cls(g, v, soa=True))
    return (ListModeDataSplitWithoutTOF if without_tof else ListModeDataSplit)(*parts)
