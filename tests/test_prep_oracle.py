"""CPU tests of the image loading and normalisation pieces: the oracle's
restatement of preprocess_image / mean_and_variance / the scaling
(oracle/prep_oracle.cpp), its Gaussian FIR, and the TIFF readers.

libcvd's convolveGaussian_fir source is not in the reference tree (only its
header is vendored), so the FIR is pinned three ways: bit for bit against a
literal float32 restatement of the header template's sliding-window code
(include/cvd/convolution.h:400-499) below, against
scipy.ndimage.gaussian_filter (the same truncation and repeated edges)
within float rounding, and on constant and linear images.  The TIFF readers
(the product's C++ one and the oracle's Python one) are pinned on bytes
assembled here by hand, on round trips in both byte orders and strip
layouts, and on the reference's own test images when present.
"""
import math
import os
import struct

import numpy as np
import pytest
import scipy.ndimage

from oracle import preporacle as po
from rfba_amd import prep

f32 = np.float32


def template_blur(I, sigma, sigmas=3.0):
    """convolveGaussian<float> (convolution.h:400-499) line by line."""
    ksize = int(math.ceil(sigmas * sigma))
    kernel = [f32(0)] * ksize
    ksum = f32(0)
    for i in range(1, ksize + 1):
        kernel[i - 1] = f32(math.exp(-i * i / (2 * sigma * sigma)))
        ksum = f32(ksum + kernel[i - 1])
    den = f32(f32(2) * ksum + f32(1))
    kernel = [f32(k / den) for k in kernel]
    factor = 1.0 / float(den)
    h, w = I.shape
    swin = 2 * ksize
    rows = [np.zeros(w, np.float32) for _ in range(swin + 1)]
    out = np.zeros((h, w), np.float32)
    o = 0

    def vsum(mid, pairs):
        ob = (mid.astype(np.float64) * factor).astype(np.float32)
        for k, (a, b) in enumerate(pairs):
            ob = (ob + (a + b) * kernel[k]).astype(np.float32)
        return ob

    for i in range(h):
        nr = rows[swin]
        inp = I[i]
        for j in range(ksize):
            s = f32(float(inp[j]) * factor)
            for k in range(ksize):
                s = f32(s + f32(f32(inp[max(j - k - 1, 0)] + inp[j + k + 1]) * kernel[k]))
            nr[j] = s
        ff = f32(factor)
        for j in range(ksize, w - ksize):
            s = f32(inp[j] * ff)
            for k in range(ksize):
                s = f32(s + f32(f32(inp[j - k - 1] + inp[j + k + 1]) * kernel[k]))
            nr[j] = s
        for j in range(w - ksize, w):
            s = f32(float(inp[j]) * factor)
            room = w - j
            for k in range(ksize):
                s = f32(s + f32(f32(inp[j - k - 1] + inp[j + min(k + 1, room - 1)]) * kernel[k]))
            nr[j] = s
        if i >= swin:
            out[o] = vsum(rows[ksize], [(rows[ksize - k - 1], rows[ksize + k + 1]) for k in range(ksize)])
            o += 1
            if i == h - 1:
                for r in range(ksize):
                    out[o] = vsum(rows[ksize + r + 1],
                                  [(rows[ksize + r - k], rows[min(ksize + r + k + 2, swin)]) for k in range(ksize)])
                    o += 1
        elif i == swin - 1:
            for r in range(ksize):
                out[o] = vsum(rows[r + 1], [(rows[max(r - k - 1, 0) + 1], rows[r + k + 2]) for k in range(ksize)])
                o += 1
        rows = rows[1:] + rows[:1]
    assert o == h
    return out


def test_gaussian_kernel_product_and_oracle():
    for sigma in (1.0, 2.5, 5.0, 7.3):
        k, f = prep.gaussian_kernel(sigma)
        ko, fo = po.kernel(sigma)
        assert np.array_equal(k, ko) and f == fo
        assert len(k) == math.ceil(3 * sigma)
        assert abs(f + 2 * k.astype(np.float64).sum() - 1) < 1e-6
        ratio = k[1:].astype(np.float64) / k[:-1]
        i = np.arange(1, len(k))
        assert np.allclose(ratio, np.exp(-((i + 1) ** 2 - i ** 2) / (2 * sigma ** 2)), rtol=1e-6)


@pytest.mark.parametrize("sigma,w,h", [(1.0, 6, 7), (1.0, 23, 19), (2.0, 12, 13), (2.0, 31, 40), (5.0, 30, 31),
                                       (5.0, 47, 36)])
def test_blur_equals_header_template_bit_for_bit(sigma, w, h):
    rng = np.random.default_rng(w * 100 + h)
    img = (rng.poisson(100, (h, w)) + rng.normal(0, 5, (h, w))).astype(np.float32)
    assert np.array_equal(po.blur(img, sigma), template_blur(img, sigma))


@pytest.mark.parametrize("sigma", [1.2, 2.5, 5.0])
def test_blur_close_to_scipy(sigma):
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1000, (64, 80)).astype(np.float32)
    ref = scipy.ndimage.gaussian_filter(img.astype(np.float64), sigma, mode="nearest", truncate=3.0)
    assert int(3.0 * sigma + 0.5) == math.ceil(3 * sigma)  # the same half-width
    # float32 sums of 2 x (2 ceil(3 sigma) + 1) terms: a few units of float rounding
    assert np.max(np.abs(po.blur(img, sigma) - ref)) < 2e-6 * np.max(np.abs(ref))


def test_highpass_constant_and_ramp():
    T, Z, Y, X = 1, 2, 40, 50
    const = np.full((T, Z, Y, X), 123.0, np.float32)
    x = np.arange(X, dtype=np.float32)
    ramp = np.broadcast_to(3.0 * x, (T, Z, Y, X)).astype(np.float32)
    series = np.concatenate([const, ramp])
    o, st = po.normalize(series, skip=0, lpf=5.0)
    s = math.sqrt(st[1])
    raw = o.astype(np.float64) * s  # undo the scaling
    assert np.all(raw[:, :, 0, :] == 0) and np.all(raw[:, :, -1, :] == 0)
    assert np.all(raw[:, :, :, 0] == 0) and np.all(raw[:, :, :, -1] == 0)
    assert np.max(np.abs(raw[0, :, 1:-1, 1:-1])) < 1e-3  # constant: |f| << 123
    # ramp: symmetric kernel, exact far from the x edges
    assert np.max(np.abs(raw[1, :, 1:-1, 16:-16])) < 1e-3
    # near an x edge the repeated edge pixels pull the blur up: f < 0 at x = 1
    assert np.all(raw[1, :, 1:-1, 1] < -1.0)


def test_mean_and_variance_and_scaling():
    rng = np.random.default_rng(5)
    v = rng.poisson(50, (3, 2, 9, 11)).astype(np.float32)
    out, st = po.normalize(v, skip=1)
    m, var = v.astype(np.float64).mean(), v.astype(np.float64).var()
    assert st[0] == np.float32(m) or abs(st[0] - m) < 1e-6 * m
    assert abs(st[1] - var) < 1e-5 * var
    assert np.array_equal(out, (v.astype(np.float64) / math.sqrt(st[1])).astype(np.float32))
    # after x / sqrt(var): mean / sqrt(var) and variance 1
    assert abs(st[2] - st[0] / math.sqrt(st[1])) < 1e-5 * abs(st[2]) and abs(st[3] - 1.0) < 1e-5
    out2, st2 = po.normalize(v, skip=1, fixed=1, sub=40.0, div=4.0)
    assert np.array_equal(out2, ((v.astype(np.float64) - 40.0) / 4.0).astype(np.float32))


def tiff_bytes(pages, bo, bits, rps):
    """A TIFF assembled by hand: header, then per page the strips and an IFD."""
    npg, h, w = pages.shape
    b = bytearray((b"II" if bo == "<" else b"MM") + struct.pack(bo + "HI", 42, 0))
    link = 4
    for p in range(npg):
        data = pages[p].astype(bo + ("u1" if bits == 8 else "u2")).tobytes()
        row = w * bits // 8
        offs, cnts = [], []
        for r0 in range(0, h, rps):
            offs.append(len(b) + r0 * row)
            cnts.append(min(rps, h - r0) * row)
        b += data
        so = len(b)
        b += struct.pack(bo + "%dI" % len(offs), *offs) + struct.pack(bo + "%dI" % len(cnts), *cnts)
        ifd = len(b)
        one = len(offs) == 1  # a single LONG is stored in the entry itself
        ents = [(256, 3, 1, w), (257, 3, 1, h), (258, 3, 1, bits), (262, 3, 1, 1),
                (273, 4, len(offs), offs[0] if one else so), (278, 3, 1, rps),
                (279, 4, len(cnts), cnts[0] if one else so + 4 * len(offs))]
        b += struct.pack(bo + "H", len(ents))
        for tag, typ, cnt, val in ents:
            b += struct.pack(bo + "HHI", tag, typ, cnt)
            b += struct.pack(bo + "HH", val, 0) if typ == 3 else struct.pack(bo + "I", val)
        b[link:link + 4] = struct.pack(bo + "I", ifd)
        link = len(b)
        b += struct.pack(bo + "I", 0)
    return bytes(b)


@pytest.mark.parametrize("bo,bits,rps", [("<", 16, 3), (">", 16, 1), ("<", 8, 5), (">", 8, 2)])
def test_tiff_readers_on_hand_built_files(tmp_path, bo, bits, rps):
    rng = np.random.default_rng(bits + rps)
    pages = rng.integers(0, 2 ** bits, (4, 5, 7)).astype(np.uint16 if bits == 16 else np.uint8)
    path = tmp_path / "hand.tif"
    path.write_bytes(tiff_bytes(pages, bo, bits, rps))
    assert np.array_equal(prep.tiff_load(path), pages)
    assert np.array_equal(po.read_tiff(path), pages)
    info = prep.tiff_info(path)
    assert (info.width, info.height, info.pages, info.bits) == (7, 5, 4, bits)


@pytest.mark.parametrize("big,rps", [(False, 0), (True, 0), (True, 3), (False, 1)])
def test_tiff_round_trip(tmp_path, big, rps):
    rng = np.random.default_rng(7)
    for dt in (np.uint8, np.uint16):
        a = rng.integers(0, np.iinfo(dt).max, (3, 11, 13)).astype(dt)
        p = tmp_path / ("rt_%s.tif" % np.dtype(dt).name)
        prep.tiff_write(p, a, big_endian=big, rows_per_strip=rps)
        assert np.array_equal(prep.tiff_load(p), a)
        assert np.array_equal(po.read_tiff(p), a)


def test_tiff_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.tif"
    p.write_bytes(b"II*\x00\x08\x00\x00\x00")  # directory offset past the end
    with pytest.raises(prep.PrepError):
        prep.tiff_load(p)
    pages = np.zeros((1, 4, 4), np.uint16)
    raw = bytearray(tiff_bytes(pages, "<", 16, 4))
    raw += b"\x00" * 8
    q = tmp_path / "comp.tif"
    # mark the page compressed (tag 259 = 5, LZW) by appending an entry is
    # awkward; flip the bits-per-sample entry to 12 instead
    i = raw.find(struct.pack("<HHI", 258, 3, 1))
    raw[i + 8:i + 10] = struct.pack("<H", 12)
    q.write_bytes(bytes(raw))
    with pytest.raises(prep.PrepError, match="bits"):
        prep.tiff_load(q)
    # a page width beyond 2^20 (a LONG value) is rejected before any size is formed
    raw = bytearray(tiff_bytes(pages, "<", 16, 4))
    i = raw.find(struct.pack("<HHI", 256, 3, 1))
    raw[i:i + 12] = struct.pack("<HHII", 256, 4, 1, 1 << 21)
    w = tmp_path / "wide.tif"
    w.write_bytes(bytes(raw))
    with pytest.raises(prep.PrepError, match="2\\^20"):
        prep.tiff_info(w)


REF_DATA = "/root/reference/Test Data"


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="reference test images not present")
@pytest.mark.parametrize("name,shape", [("Microbubule.tif", (540, 40, 40)), ("nucleus.tif", (450, 50, 50))])
def test_tiff_readers_on_reference_images(name, shape):
    path = os.path.join(REF_DATA, name)
    a = prep.tiff_load(path)
    assert a.shape == shape and a.dtype == np.uint16
    assert np.array_equal(a, po.read_tiff(path))
    assert a.max() > 0
