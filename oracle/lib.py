"""TEST INFRASTRUCTURE ONLY — ctypes binding of ``oracle/liboracle.so``.

The C library restates src/cpp/tf_custom_ops/tor.cu.cc and siddon.cu.cc
(reference repository) twice: ``*_f64`` is the FP64 oracle, ``*_f32`` the FP32
CPU port that serves as the timed CPU baseline.  All arrays here are numpy, on
the host.  LOR arrays are SoA: ``lors[c, e]`` is column ``c`` of event ``e``,
exactly the ``[ncols, N]`` tensor the reference ops receive (tor.cc:140-160,
siddon.cc:108-122).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64
_int = ctypes.c_int
_flt = ctypes.c_float


def build(force=False):
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        geo = [_i32p, _f32p, _f32p]
        for suf, outp in (("_f64", _f64p), ("_f32", _f32p)):
            getattr(L, "tor_project" + suf).argtypes = [_f32p, _i64, _i64] + geo + [
                _f32p, _flt, _int, _flt, _flt, _int, outp]
            for name in ("tor_backproject", "tor_backproject_adjoint"):
                getattr(L, name + suf).argtypes = [_f32p, _i64, _i64] + geo + [
                    _f32p, _flt, _int, _flt, _flt, _int, outp]
            getattr(L, "tor_map" + suf).argtypes = [_f32p, _i64, _i64] + geo + [
                _f32p, _flt, _int, _int, outp]
            getattr(L, "tor_count" + suf).argtypes = [_f32p, _i64, _i64] + geo + [
                _flt, _int, _int, _i64p]
            getattr(L, "siddon_project" + suf).argtypes = [_f32p, _i64, _i64] + geo + [
                _f32p, _flt, _flt, _int, outp]
            for name in ("siddon_backproject", "siddon_backproject_adjoint"):
                getattr(L, name + suf).argtypes = [_f32p, _i64, _i64] + geo + [
                    _f32p, _flt, _flt, _int, outp]
            getattr(L, "siddon_map" + suf).argtypes = [_f32p, _i64, _i64] + geo + [
                _f32p, _int, outp]
            getattr(L, "siddon_count" + suf).argtypes = [_f32p, _i64, _i64] + geo + [
                _flt, _flt, _int, _i64p]
        L.oracle_max_threads.restype = _int
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _geom(grid, center, size):
    g = np.ascontiguousarray(grid, dtype=np.int32)
    c = _f32(center)
    s = _f32(size)
    assert g.shape == (3,) and c.shape == (3,) and s.shape == (3,)
    assert np.all(g > 0)
    return g, c, s


def _p(a, t):
    return a.ctypes.data_as(t)


def _out(prec, n):
    if prec == "f64":
        return np.zeros(n, np.float64), _f64p
    if prec == "f32":
        return np.zeros(n, np.float32), _f32p
    raise ValueError(prec)


def _lors(lors, rows):
    lors = _f32(lors)
    assert lors.ndim == 2 and lors.shape[0] >= rows, lors.shape
    return lors, lors.shape[1], lors.shape[1]


def tor_project(lors, image, grid, center, size, kernel_width, geo_sigma2_flag, tof_bin,
                tof_sigma2, prec="f64", nthreads=0):
    """Reference op ``Projection`` (tor.cc:14-28): lors [10, N], image in the
    group frame laid out [gz][gy][gx].  Returns p[N]."""
    lors, n, ld = _lors(lors, 10)
    g, c, s = _geom(grid, center, size)
    img = _f32(image).reshape(-1)
    assert img.size == int(np.prod(g))
    out, t = _out(prec, n)
    getattr(lib(), "tor_project_" + prec)(
        _p(lors, _f32p), n, ld, _p(g, _i32p), _p(c, _f32p), _p(s, _f32p), _p(img, _f32p),
        kernel_width, int(bool(geo_sigma2_flag)), tof_bin, tof_sigma2, nthreads, _p(out, t))
    return out


def _tor_scatter(name, lors, values, grid, center, size, kernel_width, geo, tof_bin,
                 tof_sigma2, prec, nthreads):
    lors, n, ld = _lors(lors, 10)
    g, c, s = _geom(grid, center, size)
    v = _f32(values).reshape(-1)
    assert v.size == n
    out, t = _out(prec, int(np.prod(g)))
    getattr(lib(), name + "_" + prec)(
        _p(lors, _f32p), n, ld, _p(g, _i32p), _p(c, _f32p), _p(s, _f32p), _p(v, _f32p),
        kernel_width, int(bool(geo)), tof_bin, tof_sigma2, nthreads, _p(out, t))
    return out.reshape(g[2], g[1], g[0])


def tor_backproject(lors, lors_value, grid, center, size, kernel_width, geo_sigma2_flag,
                    tof_bin, tof_sigma2, prec="f64", nthreads=0):
    """Reference op ``Backprojection`` (tor.cc:30-47): sum_i w_ij / v_i over
    v_i > 1e-7.  Returns [gz, gy, gx]."""
    return _tor_scatter("tor_backproject", lors, lors_value, grid, center, size,
                        kernel_width, geo_sigma2_flag, tof_bin, tof_sigma2, prec, nthreads)


def tor_backproject_adjoint(lors, g, grid, center, size, kernel_width, geo_sigma2_flag,
                            tof_bin, tof_sigma2, prec="f64", nthreads=0):
    """Plain adjoint sum_i w_ij g_i (no reference counterpart)."""
    return _tor_scatter("tor_backproject_adjoint", lors, g, grid, center, size, kernel_width,
                        geo_sigma2_flag, tof_bin, tof_sigma2, prec, nthreads)


def tor_map(lors7, lors_value, grid, center, size, kernel_width, geo_sigma2_flag, prec="f64",
            nthreads=0):
    """Reference op ``Maplors`` (tor.cc:49-64): lors [7, N] = x1..z2, sigma2
    factor; no TOF, no threshold."""
    lors, n, ld = _lors(lors7, 7)
    g, c, s = _geom(grid, center, size)
    v = _f32(lors_value).reshape(-1)
    assert v.size == n
    out, t = _out(prec, int(np.prod(g)))
    getattr(lib(), "tor_map_" + prec)(
        _p(lors, _f32p), n, ld, _p(g, _i32p), _p(c, _f32p), _p(s, _f32p), _p(v, _f32p),
        kernel_width, int(bool(geo_sigma2_flag)), nthreads, _p(out, t))
    return out.reshape(g[2], g[1], g[0])


def tor_count(lors, grid, center, size, kernel_width, geo_sigma2_flag, prec="f64", nthreads=0):
    """Per-event number of non-zero (event, voxel) pairs of the TOR model."""
    lors, n, ld = _lors(lors, 10)
    g, c, s = _geom(grid, center, size)
    out = np.zeros(n, np.int64)
    getattr(lib(), "tor_count_" + prec)(
        _p(lors, _f32p), n, ld, _p(g, _i32p), _p(c, _f32p), _p(s, _f32p), kernel_width,
        int(bool(geo_sigma2_flag)), nthreads, _p(out, _i64p))
    return out


def siddon_project(lors, image, grid, center, size, tof_bin, tof_sigma2, prec="f64",
                   nthreads=0):
    """Siddon op ``Projection`` (siddon.cc:14-26): lors [7, N] = x1..z2, tof."""
    lors, n, ld = _lors(lors, 7)
    g, c, s = _geom(grid, center, size)
    img = _f32(image).reshape(-1)
    assert img.size == int(np.prod(g))
    out, t = _out(prec, n)
    getattr(lib(), "siddon_project_" + prec)(
        _p(lors, _f32p), n, ld, _p(g, _i32p), _p(c, _f32p), _p(s, _f32p), _p(img, _f32p),
        tof_bin, tof_sigma2, nthreads, _p(out, t))
    return out


def _sid_scatter(name, lors, values, grid, center, size, tof_bin, tof_sigma2, prec, nthreads):
    lors, n, ld = _lors(lors, 7)
    g, c, s = _geom(grid, center, size)
    v = _f32(values).reshape(-1)
    assert v.size == n
    out, t = _out(prec, int(np.prod(g)))
    getattr(lib(), name + "_" + prec)(
        _p(lors, _f32p), n, ld, _p(g, _i32p), _p(c, _f32p), _p(s, _f32p), _p(v, _f32p),
        tof_bin, tof_sigma2, nthreads, _p(out, t))
    return out.reshape(g[2], g[1], g[0])


def siddon_backproject(lors, lors_value, grid, center, size, tof_bin, tof_sigma2, prec="f64",
                       nthreads=0):
    """Siddon op ``Backprojection`` (siddon.cc:28-43): sum a_ij / v_i, v_i > 0."""
    return _sid_scatter("siddon_backproject", lors, lors_value, grid, center, size, tof_bin,
                        tof_sigma2, prec, nthreads)


def siddon_backproject_adjoint(lors, g, grid, center, size, tof_bin, tof_sigma2, prec="f64",
                               nthreads=0):
    return _sid_scatter("siddon_backproject_adjoint", lors, g, grid, center, size, tof_bin,
                        tof_sigma2, prec, nthreads)


def siddon_map(lors, lors_value, grid, center, size, prec="f64", nthreads=0):
    """Siddon op ``Maplors`` (siddon.cc:45-57): non-TOF, weight 1/(L*L)."""
    lors, n, ld = _lors(lors, 6)
    g, c, s = _geom(grid, center, size)
    v = _f32(lors_value).reshape(-1)
    assert v.size == n
    out, t = _out(prec, int(np.prod(g)))
    getattr(lib(), "siddon_map_" + prec)(
        _p(lors, _f32p), n, ld, _p(g, _i32p), _p(c, _f32p), _p(s, _f32p), _p(v, _f32p),
        nthreads, _p(out, t))
    return out.reshape(g[2], g[1], g[0])


def siddon_count(lors, grid, center, size, tof_bin, tof_sigma2, prec="f64", nthreads=0):
    lors, n, ld = _lors(lors, 7)
    g, c, s = _geom(grid, center, size)
    out = np.zeros(n, np.int64)
    getattr(lib(), "siddon_count_" + prec)(
        _p(lors, _f32p), n, ld, _p(g, _i32p), _p(c, _f32p), _p(s, _f32p), tof_bin, tof_sigma2,
        nthreads, _p(out, _i64p))
    return out


def max_threads():
    return lib().oracle_max_threads()
