"""TEST INFRASTRUCTURE ONLY: ctypes wrappers of the image-normalisation
oracle (oracle/prep_oracle.cpp, in oracle/liboracle.so) and an independent
TIFF reader written from the TIFF 6.0 layout (header, IFD chain, strips), as
libtiff's reading of uncompressed 8/16-bit pages serves img_load3D
(RFBA Main/storm_imagery.cc:185-254)."""
from __future__ import annotations

import ctypes
import struct
from ctypes import POINTER, c_double, c_float, c_int

import numpy as np

from . import oracle as _o

_sig = False


def lib():
    global _sig
    L = _o.lib()
    if not _sig:
        with _o.LOCK:
            if not _sig:
                fp, dp, ip = POINTER(c_float), POINTER(c_double), POINTER(c_int)
                L.oracle_prep_kernel.argtypes = [c_double, c_double, fp, c_int, ip, dp]
                L.oracle_prep_blur.argtypes = [fp, c_int, c_int, c_double, c_double, fp]
                L.oracle_prep_normalize.argtypes = [fp, c_int, c_int, c_int, c_int, c_int, c_double, c_int, c_double,
                                                    c_double, fp, dp, c_int]
                _sig = True
    return L


def _p(a, t):
    return a.ctypes.data_as(POINTER(t))


def kernel(sigma, sigmas=3.0):
    k = np.zeros(256, np.float32)
    ks, f = c_int(0), c_double(0)
    assert lib().oracle_prep_kernel(sigma, sigmas, _p(k, c_float), 256, ctypes.byref(ks), ctypes.byref(f)) == 0
    return k[:ks.value].copy(), f.value


def blur(img, sigma, sigmas=3.0):
    img = np.ascontiguousarray(img, np.float32)
    h, w = img.shape
    out = np.empty_like(img)
    lib().oracle_prep_blur(_p(img, c_float), w, h, sigma, sigmas, _p(out, c_float))
    return out


def normalize(series_tzyx, skip=0, lpf=5.0, fixed=0, sub=0.0, div=1.0, threads=None):
    """load_and_normalize_images on a float series [T,Z,Y,X] -> (out, stats)."""
    v = np.ascontiguousarray(series_tzyx, np.float32)
    T, Z, Y, X = v.shape
    out = np.empty_like(v)
    st = np.empty(4)
    lib().oracle_prep_normalize(_p(v, c_float), X, Y, Z, T, int(skip), lpf, int(fixed), sub, div, _p(out, c_float),
                                _p(st, c_double), threads or _o.threads())
    return out, st


def read_tiff(path):
    """Pixels [pages, height, width] (uint8 or uint16) of an uncompressed
    single-sample TIFF."""
    b = open(path, "rb").read()
    bo = {b"II": "<", b"MM": ">"}[b[:2]]
    assert struct.unpack(bo + "H", b[2:4])[0] == 42
    off = struct.unpack(bo + "I", b[4:8])[0]
    pages = []
    sizes = {3: 2, 4: 4}
    while off:
        n = struct.unpack(bo + "H", b[off:off + 2])[0]
        tags = {}
        for i in range(n):
            e = b[off + 2 + 12 * i: off + 14 + 12 * i]
            tag, typ, cnt = struct.unpack(bo + "HHI", e[:8])
            if typ not in sizes:
                continue
            sz = sizes[typ] * cnt
            raw = e[8:8 + sz] if sz <= 4 else b[struct.unpack(bo + "I", e[8:12])[0]:][:sz]
            tags[tag] = list(struct.unpack(bo + ("H" if typ == 3 else "I") * cnt, raw))
        w, h, bits = tags[256][0], tags[257][0], tags.get(258, [1])[0]
        assert tags.get(259, [1])[0] == 1 and tags.get(277, [1])[0] == 1 and bits in (8, 16)
        data = b"".join(b[o:o + c] for o, c in zip(tags[273], tags[279]))
        dt = np.dtype(np.uint8) if bits == 8 else np.dtype(bo + "u2")
        pages.append(np.frombuffer(data[:w * h * dt.itemsize], dt).reshape(h, w))
        off = struct.unpack(bo + "I", b[off + 2 + 12 * n: off + 6 + 12 * n])[0]
    return np.stack(pages).astype(pages[0].dtype.newbyteorder("=") if pages[0].dtype.itemsize > 1 else np.uint8)
