"""GPU test of rfba_srrf_stack_device_noalloc, the allocation-free form of
rfba_srrf_stack_device: its tables, queue counter and TRAC / TRPPM series
live in the library's static device scratch.  Called through the C ABI and
checked bit for bit against rfba_srrf_stack_device on consecutive calls,
the second with a smaller input, one whose series area limits the grid, and
a size beyond the scratch tile table (rejected on the host)."""
import ctypes
from ctypes import c_int

import numpy as np
import pytest

from rfba_amd import srrf, synth

pytestmark = pytest.mark.gpu


def conf(**kw):
    base = dict(RingRadius=0.5, Magnification=2, AxesinRing=2, MainRange=1, XRange=1, YRange=1, ZRange=1,
                Method=1, ITC=0, TRACOrder=2, RemoveCon=0, Renormalize=0, GradientSmooth=1, IntensityW=0,
                blockSizeX=16, blockSizeY=16, blockSizeZ=16)
    base.update(kw)
    return srrf.default_conf().replace(**base)


def noalloc(d_in, shape_tzyx, c, smp, d_out):
    T, Z, Y, X = shape_tzyx
    dims = (c_int * 4)(X, Y, Z, T)
    cs = smp._c()
    return srrf.lib().rfba_srrf_stack_device_noalloc(d_in.ptr, srrf.RFBA_U16, dims, ctypes.byref(c),
                                                     ctypes.byref(cs), d_out.ptr, None)


def case(shape_tzyx, n_em, seed, c, smp):
    T, Z, Y, X = shape_tzyx
    s = synth.srrf_series((X, Y, Z, T), n_em, seed)
    M = c.Magnification
    n_sr = X * Y * Z * M ** 3
    d_in = srrf.DeviceBuffer.from_array(s)
    d_a, d_b = srrf.DeviceBuffer(n_sr * 4), srrf.DeviceBuffer(n_sr * 4)
    d_a.zero()
    d_b.zero()
    print("noalloc case: T=%d %dx%dx%d M=%d method %d" % (T, X, Y, Z, M, c.Method), flush=True)
    rc = noalloc(d_in, s.shape, c, smp, d_a)
    assert rc == 0, srrf.lib().rfba_last_error()
    srrf.srrf_stack_device(d_in, srrf.RFBA_U16, s.shape, c, smp, d_b)
    a = d_a.to_array((Z * M, Y * M, X * M), np.float32)
    b = d_b.to_array((Z * M, Y * M, X * M), np.float32)
    for d in (d_in, d_a, d_b):
        d.close()
    assert np.isfinite(a).any()
    assert np.array_equal(a, b, equal_nan=True)


def test_noalloc_matches_stack_device(gpu):
    smp26 = srrf.directions(26, 0.5)
    # first call (possibly the process's first library call), then a smaller one
    case((12, 10, 18, 20), 4, 91, conf(Magnification=3, Method=3, ITC=1, IntensityW=1), smp26)
    case((6, 8, 12, 14), 3, 92, conf(Magnification=2, Method=1, IntensityW=1), srrf.directions(6, 0.5))
    # 150-frame series: the scratch series area holds fewer workgroup slots
    # than the occupancy grid, so the grid is reduced
    case((150, 12, 40, 40), 20, 93, conf(Magnification=4, Method=3, ITC=0, TRACOrder=2, IntensityW=1), smp26)
    # TRPPM through the series area as well
    case((9, 8, 14, 15), 3, 94, conf(Magnification=2, Method=2, IntensityW=1), smp26)


def test_noalloc_rejects_more_tiles_than_its_table(gpu):
    # 4096 x 4096 x 256 SR voxels need over 10^6 tiles: rejected on the host
    # before any device work (the input is never read)
    c = conf(Magnification=4, Method=1)
    d_in = srrf.DeviceBuffer(16)
    d_out = srrf.DeviceBuffer(16)
    rc = noalloc(d_in, (1, 64, 1024, 1024), c, srrf.directions(6, 0.5), d_out)
    msg = srrf.lib().rfba_last_error().decode()
    d_in.close()
    d_out.close()
    assert rc != 0 and "tile" in msg, (rc, msg)
