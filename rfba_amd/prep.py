"""Python mirror of the image loading and normalisation interface
(librfba_prep.so via ctypes).

Reference: img_load3D (RFBA Main/storm_imagery.cc:185-254), preprocess_image
(storm_imagery_no_load.cc:44-79), mean_and_variance (:87-110) and
load_and_normalize_images (storm_imagery.cc:358-416); configuration keys
preprocess.skip / lpf / fixed_scaling (multispot5.cfg:49-52).  Errors raise
PrepError; there is no CPU fallback for the normalisation.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_void_p

import numpy as np

from . import _lib


class PrepError(RuntimeError):
    pass


class TiffInfo(ctypes.Structure):
    _fields_ = [("width", c_int), ("height", c_int), ("pages", c_int), ("bits", c_int)]


class PrepConf(ctypes.Structure):
    """ABI of rfba_prep_conf (defaults: multispot5.cfg:49-52)."""
    _fields_ = [("skip", c_int), ("lpf", c_float), ("fixed_scaling", c_int), ("subtract", c_double),
                ("divide", c_double)]


def conf(skip=0, lpf=5.0, fixed_scaling=0, subtract=0.0, divide=1.0) -> PrepConf:
    return PrepConf(int(skip), float(lpf), int(fixed_scaling), float(subtract), float(divide))


U8, U16, F32 = 8, 16, 32
_TYPES = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16, np.dtype(np.float32): F32}

_sig = False


def lib():
    global _sig
    L = _lib.load("rfba_prep")
    if not _sig:
        L.rfba_tiff_info_read.argtypes = [c_char_p, POINTER(TiffInfo)]
        L.rfba_tiff_load.argtypes = [c_char_p, c_void_p, c_size_t, POINTER(TiffInfo)]
        L.rfba_tiff_write.argtypes = [c_char_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int]
        L.rfba_prep_gaussian_kernel.argtypes = [c_double, c_double, POINTER(c_float), c_int, POINTER(c_int),
                                                POINTER(c_double)]
        L.rfba_prep_normalize_device.argtypes = [c_void_p, c_int, POINTER(c_int), POINTER(PrepConf), c_void_p,
                                                 POINTER(c_double), c_void_p, POINTER(c_float)]
        L.rfba_prep_load_and_normalize.argtypes = [POINTER(c_char_p), c_int, POINTER(PrepConf), c_void_p, c_size_t,
                                                   POINTER(c_int), POINTER(c_double), c_void_p]
        L.rfba_prep_malloc.argtypes = [POINTER(c_void_p), c_size_t]
        L.rfba_prep_free.argtypes = [c_void_p]
        L.rfba_prep_memcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
        L.rfba_prep_last_error.restype = c_char_p
        for f in ("rfba_tiff_info_read", "rfba_tiff_load", "rfba_tiff_write", "rfba_prep_gaussian_kernel",
                  "rfba_prep_normalize_device", "rfba_prep_load_and_normalize", "rfba_prep_malloc", "rfba_prep_free",
                  "rfba_prep_memcpy"):
            getattr(L, f).restype = c_int
        _sig = True
    return L


def _check(rc):
    if rc != 0:
        raise PrepError("rfba_prep error %d: %s" % (rc, lib().rfba_prep_last_error().decode(errors="replace")))


class Device:
    """A raw device allocation (freed on close/del)."""

    def __init__(self, nbytes):
        self.ptr = c_void_p()
        self.nbytes = int(nbytes)
        _check(lib().rfba_prep_malloc(ctypes.byref(self.ptr), max(self.nbytes, 1)))

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        _check(lib().rfba_prep_memcpy(d.ptr, a.ctypes.data_as(c_void_p), a.nbytes, 1))
        return d

    def to_array(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().rfba_prep_memcpy(out.ctypes.data_as(c_void_p), self.ptr, out.nbytes, 2))
        return out

    def close(self):
        if self.ptr:
            lib().rfba_prep_free(self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tiff_info(path) -> TiffInfo:
    info = TiffInfo()
    _check(lib().rfba_tiff_info_read(str(path).encode(), ctypes.byref(info)))
    return info


def tiff_load(path) -> np.ndarray:
    """img_load3D's pixels: [pages, height, width] uint8 or uint16."""
    info = tiff_info(path)
    out = np.empty((info.pages, info.height, info.width), np.uint8 if info.bits == 8 else np.uint16)
    _check(lib().rfba_tiff_load(str(path).encode(), out.ctypes.data_as(c_void_p), out.nbytes, None))
    return out


def tiff_write(path, stack, big_endian=False, rows_per_strip=0):
    """Uncompressed multi-page TIFF of a [pages, height, width] uint8/uint16 stack."""
    a = np.ascontiguousarray(stack)
    assert a.dtype in (np.uint8, np.uint16) and a.ndim == 3
    _check(lib().rfba_tiff_write(str(path).encode(), a.ctypes.data_as(c_void_p), a.shape[2], a.shape[1], a.shape[0],
                                 8 * a.itemsize, int(big_endian), int(rows_per_strip)))


def gaussian_kernel(sigma, sigmas=3.0):
    """(kernel[ksize] float32, centre factor) of libcvd's convolveGaussian FIR."""
    k = np.zeros(64, np.float32)
    ks, f = c_int(0), c_double(0)
    _check(lib().rfba_prep_gaussian_kernel(sigma, sigmas, k.ctypes.data_as(POINTER(c_float)), 64, ctypes.byref(ks),
                                           ctypes.byref(f)))
    return k[:ks.value].copy(), f.value


def normalize_device(d_raw, dtype, shape_tzyx, c: PrepConf, d_out):
    """Device series -> device float series; returns (stats[4], kernel ms)."""
    T, Z, Y, X = shape_tzyx
    dims = (c_int * 4)(X, Y, Z, T)
    st = (c_double * 4)()
    ms = c_float(0)
    _check(lib().rfba_prep_normalize_device(d_raw.ptr, _TYPES[np.dtype(dtype)], dims, ctypes.byref(c), d_out.ptr, st,
                                            None, ctypes.byref(ms)))
    return np.array(st[:]), ms.value


def normalize(series_tzyx, c: PrepConf | None = None):
    """Host series [T,Z,Y,X] (uint8/uint16/float32) -> (float32 series, stats)."""
    c = c or conf()
    a = np.ascontiguousarray(series_tzyx)
    d_in = Device.from_array(a)
    d_out = Device(a.size * 4)
    try:
        stats, _ = normalize_device(d_in, a.dtype, a.shape, c, d_out)
        return d_out.to_array(a.shape, np.float32), stats
    finally:
        d_in.close()
        d_out.close()


def load_and_normalize(paths, c: PrepConf | None = None):
    """load_and_normalize_images: TIFF stacks (one per frame) -> (float32
    series [T,Z,Y,X], stats)."""
    c = c or conf()
    info = tiff_info(paths[0])
    n = len(paths)
    shape = (n, info.pages, info.height, info.width)
    d_out = Device(int(np.prod(shape)) * 4)
    arr = (c_char_p * n)(*[str(p).encode() for p in paths])
    dims = (c_int * 4)()
    st = (c_double * 4)()
    try:
        _check(lib().rfba_prep_load_and_normalize(arr, n, ctypes.byref(c), d_out.ptr, int(np.prod(shape)), dims, st,
                                                  None))
        return d_out.to_array(shape, np.float32), np.array(st[:])
    finally:
        d_out.close()
