"""SRRF-guided spot placement (multispot5_place_methods.cc:293-341).

The oracle (oracle/srrf_oracle.cpp) is pinned here on the CPU: the SR pixel
expansion and volume index (a single hot SR cell receives every spot, each
within half an SR pixel of that cell), the acceptance rule (two hot cells
with values 1 and 3 receive spots in that ratio, binomial bounds), and the
negative-maximum rule (no spots).  The GPU path is compared with it
exactly."""
import numpy as np
import pytest

from oracle import oracle
from rfba_amd import srrf


def sr_volume(dims_xyz, mag):
    X, Y, Z = dims_xyz
    return np.zeros((Z * mag, Y * mag, X * mag), np.float32)


def test_single_hot_cell_receives_every_spot():
    mag = 3
    sr = sr_volume((4, 5, 3), mag)
    cell = (5, 7, 4)  # SR (z, y, x)
    sr[cell] = 2.0
    pixels = [[x, y, z] for z in range(3) for y in range(5) for x in range(4)]
    sp = oracle.place_spots(sr, mag, pixels, 50, 7.0, 0.3, seed=11)
    assert len(sp) == 50 and np.all(sp[:, 0] == 7.0) and np.all(sp[:, 1] == np.float64(0.3))
    # SR coordinate = low-res * mag, within half an SR pixel of the cell (x, y, z)
    for k, c in ((2, cell[2]), (3, cell[1]), (4, cell[0])):
        s = sp[:, k] * mag
        assert np.all(s >= c - 0.5 - 1e-5) and np.all(s <= c + 0.5 + 1e-5)


def test_acceptance_proportional_to_value():
    mag = 2
    sr = sr_volume((3, 3, 3), mag)
    sr[1, 2, 3] = 1.0
    sr[4, 4, 0] = 3.0
    pixels = [[x, y, z] for z in range(3) for y in range(3) for x in range(3)]
    sp = oracle.place_spots(sr, mag, pixels, 4000, 1.0, 1.0, seed=5)
    near_b = np.abs(sp[:, 4] * mag - 4) <= 0.51
    frac = near_b.mean()
    # expected 3/4; binomial sd = sqrt(.75 * .25 / 4000) = 0.0068
    assert abs(frac - 0.75) < 0.03, frac


def test_negative_maximum_places_nothing():
    mag = 2
    sr = np.full((4, 4, 4), -1.0, np.float32)
    assert len(oracle.place_spots(sr, mag, [[0, 0, 0], [1, 1, 1]], 10, 1.0, 1.0, seed=1)) == 0


def test_region_subset_only():
    # cells outside the region are never chosen
    mag = 2
    sr = sr_volume((4, 4, 4), mag)
    sr[0, 0, 0] = 100.0   # outside the region below
    sr[5, 5, 5] = 1.0     # inside pixel (2, 2, 2)
    sp = oracle.place_spots(sr, mag, [[2, 2, 2], [3, 3, 3]], 20, 1.0, 1.0, seed=2)
    assert np.all(np.abs(sp[:, 2:] * mag - 5) <= 0.51)


@pytest.mark.gpu
def test_gpu_placement_matches_oracle(gpu):
    from rfba_amd import synth
    s = synth.srrf_series((20, 18, 10, 12), 4, 31)
    c = srrf.default_conf().replace(Magnification=2, Method=1, IntensityW=1, GradientSmooth=1)
    vol = srrf.srrf_stack(s, c, srrf.directions(6, 0.5))
    vol = np.where(np.isnan(vol) | (vol < 0), 0, vol).astype(np.float32)
    d = srrf.DeviceBuffer.from_array(vol)
    pixels = [[x, y, z] for z in range(2, 8) for y in range(3, 15) for x in range(2, 17)]
    got, used = srrf.place_spots(d, (20, 18, 10), 2, pixels, 300, 5.0, 1.3, seed=9)
    d.close()
    ref = oracle.place_spots(vol, 2, pixels, 300, 5.0, 1.3, seed=9)
    assert got.shape == (300, 5) and used >= 3 * 300
    assert np.array_equal(got, ref)
