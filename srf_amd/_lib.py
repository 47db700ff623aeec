"""ctypes binding of libsrfx.so (the HIP kernels behind include/srfx.h).

The library is built in-tree by ``python -m srf_amd.build`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there is
no CPU fallback for any op of this package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsrfx.so")

SRFX_OK = 0
SRFX_EINVAL = 1
SRFX_ELAUNCH = 2
SRFX_EUNSUPPORTED = 3
SRFX_ENOMEM = 4
SRFX_ETOOLARGE = 5
# floats of the static scratch of srfx_tor_osem_subset_noalloc (slabs x events)
NOALLOC_SCRATCH_FLOATS = 1 << 22
SRFX_ACCUMULATE = 1
SRFX_FULL_PATCH = 2
SRFX_LPE_SHIFT = 4
SRFX_BLOCK_SHIFT = 6
_LPE_CODE = {1: 1, 4: 2, 16: 3}
_BLOCK_CODE = {64: 1, 128: 2, 256: 3}
SRFX_SLAB_SHIFT = 8
SRFX_TILED_PROJECTION = 1 << 10
SRFX_OSEM_PROJECT_ONLY = 1 << 11
SRFX_OSEM_SCATTER_ONLY = 1 << 12
SRFX_IMAGE_OP_FRAME = 1 << 13
_SLAB_CODE = {None: 0, "xcd8": 1, "blocked8": 2, "xcd16": 3}


def schedule_flags(schedule):
    """Flags for a TOR launch schedule or None: (lanes_per_event, block) or
    (lanes_per_event, block, slabs) with slabs one of None, "xcd8",
    "blocked8", "xcd16" (slab schedules take lanes_per_event 1)."""
    if schedule is None:
        return 0
    lpe, block = schedule[0], schedule[1]
    slab = schedule[2] if len(schedule) > 2 else None
    if lpe not in _LPE_CODE or block not in _BLOCK_CODE or slab not in _SLAB_CODE:
        raise ValueError(f"unsupported schedule {schedule}: lanes per event in (1, 4, 16), "
                         "block in (64, 128, 256), slabs in (None, 'xcd8', 'blocked8', "
                         "'xcd16')")
    if slab is not None and lpe != 1:
        raise ValueError("slab schedules run one lane per event")
    return ((_LPE_CODE[lpe] << SRFX_LPE_SHIFT) | (_BLOCK_CODE[block] << SRFX_BLOCK_SHIFT)
            | (_SLAB_CODE[slab] << SRFX_SLAB_SHIFT))


class SrfxGrid(ctypes.Structure):
    """srfx_grid: shape (gx, gy, gz), center, size (full extent, mm) and the
    element strides of one step along the op's x, y, z (all 0 = contiguous
    [gz][gy][gx])."""
    _fields_ = [("shape", ctypes.c_int32 * 3), ("center", ctypes.c_float * 3),
                ("size", ctypes.c_float * 3), ("stride", ctypes.c_int64 * 3)]


class SrfxTorModel(ctypes.Structure):
    _fields_ = [("kernel_width", ctypes.c_float), ("geo_sigma2_flag", ctypes.c_int32),
                ("tof_bin", ctypes.c_float), ("tof_sigma2", ctypes.c_float)]


class SrfxTile(ctypes.Structure):
    """srfx_tile: tile of `slices` x `wx` x `wy` op-frame voxels, at most
    `chunk` events per workgroup, `block` threads per workgroup; `projection`
    keeps the tables of the tiled projection."""
    _fields_ = [("slices", ctypes.c_int32), ("wx", ctypes.c_int32), ("wy", ctypes.c_int32),
                ("chunk", ctypes.c_int32), ("block", ctypes.c_int32),
                ("projection", ctypes.c_int32)]


class SrfxSubsetGroup(ctypes.Structure):
    """srfx_subset_group: one LOR array (SoA rows, n events, row stride ld) of
    a subset-plan build, its frame, its tile and the bounds of its plans."""
    _fields_ = [("lors", ctypes.c_void_p), ("n", ctypes.c_int64), ("ld", ctypes.c_int64),
                ("grid", SrfxGrid), ("tile", SrfxTile), ("nplans", ctypes.c_int32),
                ("bounds", ctypes.POINTER(ctypes.c_int64))]


class SrfxSiddonTile(ctypes.Structure):
    """srfx_siddon_tile: image tiles of `tx` x `ty` x `tz` voxels, at most
    `chunk` runs per workgroup, `block` threads per workgroup; `sort` keeps a
    direction/position-sorted copy of the LORs."""
    _fields_ = [("tx", ctypes.c_int32), ("ty", ctypes.c_int32), ("tz", ctypes.c_int32),
                ("chunk", ctypes.c_int32), ("block", ctypes.c_int32), ("sort", ctypes.c_int32)]


class SrfxError(RuntimeError):
    def __init__(self, name, status, detail):
        super().__init__(f"{name} failed with status {status} ({detail})")
        self.status = status


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32
_F = ctypes.c_float
_I32 = ctypes.c_int32

# name -> argtypes (all return int)
_PROTOS = {
    "srfx_tor_project": [_P, _I64, _I64, _P, SrfxGrid, SrfxTorModel, _P, _U32, _P],
    "srfx_tor_backproject": [_P, _I64, _I64, _P, SrfxGrid, SrfxTorModel, _P, _U32, _P],
    "srfx_tor_backproject_adjoint": [_P, _I64, _I64, _P, SrfxGrid, SrfxTorModel, _P, _U32, _P],
    "srfx_tor_maplors": [_P, _I64, _I64, _P, SrfxGrid, _F, _I32, _P, _U32, _P],
    "srfx_tor_osem_subset": [_P, _I64, _I64, _P, SrfxGrid, SrfxGrid, SrfxTorModel, _P, _U32, _P],
    "srfx_tor_count_nnz": [_P, _I64, _I64, SrfxGrid, SrfxTorModel, _P, _P, _P],
    "srfx_tor_osem_subset_noalloc": [_P, _I64, _I64, _P, SrfxGrid, SrfxGrid, SrfxTorModel, _P,
                                     _U32, _P],
    "srfx_tor_patch_info": [SrfxGrid, _F, ctypes.POINTER(_I32), ctypes.POINTER(_I32)],
    "srfx_siddon_project": [_P, _I64, _I64, _P, SrfxGrid, _F, _F, _P, _U32, _P],
    "srfx_siddon_backproject": [_P, _I64, _I64, _P, SrfxGrid, _F, _F, _P, _U32, _P],
    "srfx_siddon_backproject_adjoint": [_P, _I64, _I64, _P, SrfxGrid, _F, _F, _P, _U32, _P],
    "srfx_siddon_maplors": [_P, _I64, _I64, _P, SrfxGrid, _P, _U32, _P],
    "srfx_siddon_osem_subset": [_P, _I64, _I64, _P, SrfxGrid, _F, _F, _P, _U32, _P],
    "srfx_siddon_count_nnz": [_P, _I64, _I64, SrfxGrid, _F, _F, _P, _P, _P],
    "srfx_mlem_update": [_P, _P, _P, _I64, _P],
    "srfx_frame_copy": [_P, SrfxGrid, _P, _P],
    "srfx_tor_plan_create": [_P, _I64, _I64, _I32, SrfxGrid, SrfxTorModel, SrfxTile,
                             ctypes.POINTER(ctypes.c_void_p), _P],
    "srfx_tor_plan_create_subsets": [_P, _I64, _I64, _I32, SrfxGrid, SrfxTorModel, SrfxTile,
                                     _I32, ctypes.POINTER(_I64), ctypes.POINTER(ctypes.c_void_p),
                                     _P],
    "srfx_tor_plan_create_subset_groups": [ctypes.POINTER(SrfxSubsetGroup), _I32, _I32,
                                           SrfxTorModel, ctypes.POINTER(ctypes.c_void_p), _P],
    "srfx_tor_plan_info": [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I32)],
    "srfx_tor_plan_walks": [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32)],
    "srfx_tor_plan_destroy": [_P],
    "srfx_tor_plan_project": [_P, _P, _P, _U32, _P],
    "srfx_tor_plan_backproject": [_P, _P, _P, _U32, _P],
    "srfx_tor_plan_backproject_adjoint": [_P, _P, _P, _U32, _P],
    "srfx_tor_plan_maplors": [_P, _P, _P, _U32, _P],
    "srfx_tor_plan_osem_subset": [_P, _P, SrfxGrid, _P, _U32, _P],
    "srfx_tor_plan_scales": [_P, _P, _P],
    "srfx_siddon_plan_create": [_P, _I64, _I64, SrfxGrid, _F, _F, SrfxSiddonTile,
                                ctypes.POINTER(ctypes.c_void_p), _P],
    "srfx_siddon_plan_create_subsets": [_P, _I64, _I64, SrfxGrid, _F, _F, SrfxSiddonTile, _I32,
                                        ctypes.POINTER(_I64), ctypes.POINTER(ctypes.c_void_p),
                                        _P],
    "srfx_siddon_plan_info": [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I32)],
    "srfx_siddon_plan_destroy": [_P],
    "srfx_siddon_plan_backproject": [_P, _P, _P, _U32, _P],
    "srfx_siddon_plan_backproject_adjoint": [_P, _P, _P, _U32, _P],
    "srfx_siddon_plan_osem_subset": [_P, _P, _P, _P, _U32, _P],
    "srfx_siddon_plan_projection": [_P, _P, _P],
}

_lib = None


def load():
    """Load libsrfx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m srf_amd.build` "
            "(this package has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _PROTOS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.srfx_strerror.argtypes = [ctypes.c_int]
    lib.srfx_strerror.restype = ctypes.c_char_p
    _lib = lib
    return lib


# bound in srf_amd/batch.py (they take the batch structs)
_BATCH = ["srfx_tor_batch_create", "srfx_tor_batch_subset", "srfx_tor_batch_set_schedule",
          "srfx_tor_batch_destroy"]
_PSF = ["srfx_psf_apply"]  # bound in psf.py


def exported_symbols():
    return list(_PROTOS) + ["srfx_strerror"] + _BATCH + _PSF


def call(name, *args):
    lib = load()
    status = getattr(lib, name)(*args)
    if status != SRFX_OK:
        raise SrfxError(name, status, lib.srfx_strerror(status).decode())
    return status


def make_grid(shape, center, size, stride=(0, 0, 0)):
    g = SrfxGrid()
    for k in range(3):
        g.shape[k] = int(shape[k])
        g.center[k] = float(center[k])
        g.size[k] = float(size[k])
        g.stride[k] = int(stride[k])
    return g


def make_tor_model(kernel_width, geo_sigma2_flag, tof_bin, tof_sigma2):
    m = SrfxTorModel()
    m.kernel_width = float(kernel_width)
    m.geo_sigma2_flag = 1 if geo_sigma2_flag else 0
    m.tof_bin = float(tof_bin)
    m.tof_sigma2 = float(tof_sigma2)
    return m


def tor_patch_info(grid, size, kernel_width):
    """(P, box) for a TOR op frame (grid (gx, gy, gz), full size in mm): the
    reference's patch width P (tor.cu.cc:375) and the candidate box the
    kernels run (3, 4, or 0 for the generic chord walk).  Host only."""
    p, b = _I32(), _I32()
    call("srfx_tor_patch_info", make_grid(grid, (0, 0, 0), size), float(kernel_width),
         ctypes.byref(p), ctypes.byref(b))
    return int(p.value), int(b.value)
