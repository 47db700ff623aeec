"""The C ABI without a GPU: libsrfx.so loads, exports every entry point that
include/srfx.h declares, and rejects invalid arguments before touching the
device (no compute call is made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "srfx.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(srfx_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_entry_points():
    names = declared()
    for n in ("srfx_tor_project", "srfx_tor_backproject", "srfx_tor_maplors",
              "srfx_siddon_project", "srfx_siddon_backproject", "srfx_siddon_maplors",
              "srfx_tor_osem_subset", "srfx_mlem_update", "srfx_strerror"):
        assert n in names


def test_library_exports_all_declared_symbols():
    from srf_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
    assert sorted(_lib.exported_symbols()) == declared()


def _grid(shape=(4, 4, 4), size=(8.0, 8.0, 8.0)):
    from srf_amd import _lib
    return _lib.make_grid(shape, (0, 0, 0), size)


def test_invalid_arguments_are_rejected():
    from srf_amd import _lib
    lib = _lib.load()
    m = _lib.make_tor_model(4.0, 0, 6.0, 1000.0)
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    # n < 0
    assert lib.srfx_tor_project(fake, -1, 0, fake, _grid(), m, fake, 0, None) == _lib.SRFX_EINVAL
    # ld < n
    assert lib.srfx_tor_project(fake, 10, 5, fake, _grid(), m, fake, 0, None) == _lib.SRFX_EINVAL
    # null LORs with n > 0
    assert lib.srfx_tor_project(None, 10, 10, fake, _grid(), m, fake, 0, None) == _lib.SRFX_EINVAL
    # non-positive grid / size
    assert lib.srfx_tor_project(fake, 10, 10, fake, _grid((0, 4, 4)), m, fake, 0,
                                None) == _lib.SRFX_EINVAL
    assert lib.srfx_tor_project(fake, 10, 10, fake, _grid(size=(8.0, -1.0, 8.0)), m, fake, 0,
                                None) == _lib.SRFX_EINVAL
    # kernel width <= 0
    bad = _lib.make_tor_model(0.0, 0, 6.0, 1000.0)
    assert lib.srfx_tor_project(fake, 10, 10, fake, _grid(), bad, fake, 0,
                                None) == _lib.SRFX_EINVAL
    # TOR patch wider than supported (tiny voxels, wide kernel)
    wide = _lib.make_tor_model(40.0, 0, 6.0, 1000.0)
    assert lib.srfx_tor_project(fake, 10, 10, fake, _grid(size=(4.0, 4.0, 4.0)), wide, fake, 0,
                                None) == _lib.SRFX_EUNSUPPORTED
    # TOF on with a non-positive tof_sigma2
    neg = _lib.make_tor_model(4.0, 0, 6.0, -5.0)
    assert lib.srfx_tor_backproject(fake, 10, 10, fake, _grid(), neg, fake, 0,
                                    None) == _lib.SRFX_EINVAL
    # Siddon: ld < n, null output
    assert lib.srfx_siddon_project(fake, 10, 3, fake, _grid(), 6.0, 1e6, fake, 0,
                                   None) == _lib.SRFX_EINVAL
    assert lib.srfx_siddon_backproject(fake, 10, 10, fake, _grid(), 6.0, 1e6, None, 0,
                                       None) == _lib.SRFX_EINVAL
    # OSEM: image and bp grids describing different frames
    assert lib.srfx_tor_osem_subset(fake, 10, 10, fake, _grid(), _grid((4, 4, 5)), m, fake, 0,
                                    None) == _lib.SRFX_EINVAL
    # update: negative length
    assert lib.srfx_mlem_update(fake, fake, fake, -3, None) == _lib.SRFX_EINVAL
    # PSF blur: negative rows, blur in place, nz beyond the LDS rows
    from srf_amd import psf
    psf._bind()
    other = ctypes.c_void_p(0x2000)
    assert lib.srfx_psf_apply(fake, fake, fake, -1, fake, 4, 3, 3, fake, other,
                              None) == _lib.SRFX_EINVAL
    assert lib.srfx_psf_apply(fake, fake, fake, 8, fake, 4, 3, 3, fake, fake,
                              None) == _lib.SRFX_EINVAL
    assert lib.srfx_psf_apply(fake, fake, fake, 8, fake, 4, -1, 3, fake, other,
                              None) == _lib.SRFX_EINVAL
    assert lib.srfx_psf_apply(fake, fake, fake, 8, fake, 2049, 0, 0, fake, other,
                              None) == _lib.SRFX_EUNSUPPORTED
    # frame copy: in place, null output
    assert lib.srfx_frame_copy(fake, _grid(), fake, None) == _lib.SRFX_EINVAL
    assert lib.srfx_frame_copy(fake, _grid(), None, None) == _lib.SRFX_EINVAL
    # plan queries: null plan
    assert lib.srfx_tor_plan_walks(None, None, None) == _lib.SRFX_EINVAL
    # grouped subset plans: negative count, null groups, null output, a
    # group with the tiled projection, decreasing bounds, bounds past n
    hs = (ctypes.c_void_p * 4)()
    bs = (ctypes.c_int64 * 3)(0, 6, 10)
    tile = _lib.SrfxTile(10, 12, 32, 2048, 256, 0)

    def group(bounds=bs, n=10, t=tile):
        return (_lib.SrfxSubsetGroup * 1)(_lib.SrfxSubsetGroup(fake.value, n, n, _grid(), t, 2,
                                                               bounds))
    create = lib.srfx_tor_plan_create_subset_groups
    assert create(group(), -1, 10, m, hs, None) == _lib.SRFX_EINVAL
    assert create(None, 1, 10, m, hs, None) == _lib.SRFX_EINVAL
    assert create(group(), 1, 10, m, None, None) == _lib.SRFX_EINVAL
    assert create(group(t=_lib.SrfxTile(10, 12, 32, 2048, 256, 1)), 1, 10, m, hs,
                  None) == _lib.SRFX_EINVAL
    assert create(group(bounds=(ctypes.c_int64 * 3)(0, 6, 4)), 1, 10, m, hs,
                  None) == _lib.SRFX_EINVAL
    assert create(group(n=8), 1, 10, m, hs, None) == _lib.SRFX_EINVAL
    assert lib.srfx_strerror(_lib.SRFX_EINVAL) == b"invalid argument"


def test_python_wrappers_refuse_host_tensors():
    import torch
    from srf_amd import ops
    lors = torch.zeros(10, 4)
    img = torch.zeros(4, 4, 4)
    with pytest.raises(ValueError, match="GPU"):
        ops.projection(lors, img, (4, 4, 4), (0, 0, 0), (8, 8, 8), 4.0)
