"""GPU parity of the device normalisation (librfba_prep.so) against the
oracle's restatement of load_and_normalize_images (storm_imagery.cc:358-416):
the high-pass per z-slice with libcvd's FIR, mean_and_variance and the
scaling.  Bar: the float outputs and the float statistics are identical.
The statistics are double sums in another order on the device, rounded to
float as the reference rounds them (mean_and_variance returns floats), so
a difference could only come from a sum within 1e-16 of a float rounding
boundary.
"""
import numpy as np
import pytest

from oracle import preporacle as po
from rfba_amd import prep

pytestmark = pytest.mark.gpu


def series(shape, dtype, seed):
    rng = np.random.default_rng(seed)
    v = rng.poisson(120, shape) + rng.integers(0, 40, shape)
    if dtype == np.float32:
        return (v + rng.normal(0, 3, shape)).astype(np.float32)
    return np.clip(v, 0, np.iinfo(dtype).max).astype(dtype)


CASES = [
    ((3, 2, 48, 64), np.uint16, dict(lpf=5.0)),               # partial tiles in y
    ((2, 3, 31, 30), np.uint16, dict(lpf=5.0)),               # the smallest slice the blur takes
    ((2, 2, 70, 130), np.uint8, dict(lpf=5.0)),               # several tiles per slice
    ((2, 2, 40, 33), np.float32, dict(lpf=2.5)),              # float input, 8-tap half-width
    ((1, 4, 17, 9), np.uint16, dict(lpf=1.0)),                # 3-tap half-width
    ((1, 2, 80, 75), np.uint16, dict(lpf=12.0)),              # half-width 36: an 84 KB LDS window
    ((2, 2, 5, 7), np.uint16, dict(skip=1)),                  # no high-pass
    ((2, 2, 9, 11), np.uint16, dict(skip=1, fixed_scaling=1, subtract=100.0, divide=7.5)),
    ((2, 2, 112, 208), np.uint16, dict(lpf=5.0)),             # interior tiles: 4-value window loads
    ((1, 2, 100, 170), np.uint16, dict(lpf=5.0)),             # rows not 8-byte aligned: element loads
]


@pytest.mark.parametrize("shape,dtype,kw", CASES)
def test_normalize_matches_oracle(gpu, shape, dtype, kw):
    v = series(shape, dtype, sum(shape))
    c = prep.conf(**kw)
    out, st = prep.normalize(v, c)
    oo, ost = po.normalize(v.astype(np.float32), skip=c.skip, lpf=c.lpf, fixed=c.fixed_scaling, sub=c.subtract,
                           div=c.divide)
    assert np.array_equal(st, ost), (st, ost)
    assert np.array_equal(out, oo)
    assert abs(st[3] - (1.0 if not c.fixed_scaling else st[3])) < 1e-5


def test_load_and_normalize_from_tiff_files(gpu, tmp_path):
    rng = np.random.default_rng(9)
    paths = []
    stacks = []
    for t in range(4):
        s = np.clip(rng.poisson(200, (5, 40, 36)) + 50 * t, 0, 65535).astype(np.uint16)
        p = tmp_path / ("frame%02d.tif" % t)
        prep.tiff_write(p, s, big_endian=bool(t & 1), rows_per_strip=7)
        paths.append(p)
        stacks.append(s)
    out, st = prep.load_and_normalize(paths)
    oo, ost = po.normalize(np.stack(stacks).astype(np.float32), lpf=5.0)
    assert out.shape == (4, 5, 40, 36)
    assert np.array_equal(st, ost) and np.array_equal(out, oo)


def test_normalize_rejects_what_the_reference_cannot_do(gpu):
    with pytest.raises(prep.PrepError, match="smaller"):
        prep.normalize(series((1, 1, 20, 64), np.uint16, 1), prep.conf(lpf=5.0))
    with pytest.raises(prep.PrepError, match="skip"):
        prep.normalize(series((1, 1, 40, 64), np.uint16, 1), prep.conf(fixed_scaling=1, divide=2.0))
