"""CPU tests that pin the SRRF oracle (oracle/srrf_oracle.cpp).

No reference golden vectors exist for this path (the reference ships no
tests and its result TIFFs are absent), so each oracle function is pinned by
mathematics that a plausible transcription mistake would break: textbook
Catmull-Rom values (Keys 1981, a = -1/2), partition of unity and exact
reproduction of linear/quadratic fields, closed-form gradients of ramps,
sphere geometry of the ring, exact invariants of radiality (power-of-two
scaling, integer offsets, constant input -> NaN), a radially symmetric
blob, and hand-derived temporal reductions including the reference's
in-place ITC binning quirks (kernel.cu:509-550).
"""
import numpy as np
import pytest

from oracle import oracle
from rfba_amd import srrf


def conf(**kw):
    c = srrf.SRRFConf()
    base = dict(RingRadius=0.5, Magnification=2, AxesinRing=2, MainRange=1, XRange=1, YRange=1, ZRange=1,
                Method=1, ITC=0, TRACOrder=2, RemoveCon=0, Renormalize=0, GradientSmooth=1, IntensityW=0,
                blockSizeX=16, blockSizeY=16, blockSizeZ=16)
    base.update(kw)
    for k, v in base.items():
        setattr(c, k, v)
    return c


# ------------------------------------------------------------------ cubic
def keys(x, a=-0.5):
    """Keys (1981) cubic convolution kernel, a = -1/2 (Catmull-Rom)."""
    x = np.abs(np.asarray(x, np.float64))
    return np.where(x <= 1, (a + 2) * x ** 3 - (a + 3) * x ** 2 + 1,
                    np.where(x < 2, a * x ** 3 - 5 * a * x ** 2 + 8 * a * x - 4 * a, 0.0))


def test_cubic_textbook_values():
    xs = np.linspace(-2.5, 2.5, 401)
    np.testing.assert_allclose(oracle.cubic(xs, f64=True), keys(xs), atol=1e-14)
    # FP32: the |x| in [1,2) branch cancels terms of size ~4 (ulp 4.8e-7)
    np.testing.assert_allclose(oracle.cubic(xs), keys(xs), atol=1e-6)
    assert oracle.cubic([0.0])[0] == 1.0 and oracle.cubic([1.0])[0] == 0.0 and oracle.cubic([2.0])[0] == 0.0
    assert oracle.cubic([0.5], f64=True)[0] == 0.5625 and oracle.cubic([1.5], f64=True)[0] == -0.0625


def test_cubic_partition_of_unity():
    for x in np.linspace(1.5, 3.5, 57):
        u0 = np.floor(x - 0.5)
        w = oracle.cubic([x - (u0 - 1 + i + 0.5) for i in range(4)], f64=True)
        assert abs(w.sum() - 1.0) < 1e-14
        # first moment: sum_i w_i * centre_i == x (linear reproduction)
        c = np.array([u0 - 1 + i + 0.5 for i in range(4)])
        assert abs((w * c).sum() - x) < 1e-12


# ------------------------------------------------------------------ interpolation
def _field(f, shape):
    T, Z, Y, X = shape
    t, z, y, x = np.meshgrid(np.arange(T), np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    return f(x + 0.5, y + 0.5, z + 0.5, t)


@pytest.mark.parametrize("mag", [1, 2, 3, 4])
def test_interp_reproduces_quadratics(mag):
    shape = (2, 9, 10, 11)
    f = lambda x, y, z, t: 3 + 2 * x - y + 0.5 * z + 0.25 * x * x - 0.125 * y * z + t
    arr = _field(f, shape).astype(np.float64)
    rng = np.random.default_rng(0)
    # interior SR points (valid region [1.5, dim-1.5] in low-res units)
    pts = np.stack([rng.uniform(1.5, 11 - 1.5, 200), rng.uniform(1.5, 10 - 1.5, 200),
                    rng.uniform(1.5, 9 - 1.5, 200)], 1) * mag
    got = oracle.interp(arr, mag, 1, pts, f64=True)
    lx, ly, lz = (pts / mag).T
    np.testing.assert_allclose(got, f(lx, ly, lz, 1), rtol=1e-12, atol=1e-11)
    got32 = oracle.interp(arr, mag, 1, pts)
    np.testing.assert_allclose(got32, f(lx, ly, lz, 1), rtol=2e-5, atol=2e-4)


def test_interp_zero_outside_valid_region():
    arr = np.ones((1, 8, 8, 8))
    pts = np.array([[1.49, 4, 4], [4, 1.49, 4], [4, 4, 1.49], [6.51, 4, 4], [1.5, 1.5, 1.5], [6.5, 6.5, 6.5]])
    got = oracle.interp(arr, 1, 0, pts, f64=True)
    assert list(got[:4]) == [0, 0, 0, 0]
    np.testing.assert_allclose(got[4:], 1.0, atol=1e-14)


def test_interp_brute_force_small_grid():
    # brute force: direct double sum over the 64 taps with Keys weights
    rng = np.random.default_rng(1)
    arr = rng.normal(size=(1, 6, 6, 6))
    for _ in range(50):
        p = rng.uniform(1.5, 4.5, 3)
        u0 = np.floor(p - 0.5).astype(int)
        s = 0.0
        for k in range(4):
            for j in range(4):
                for i in range(4):
                    u, v, w = u0[0] - 1 + i, u0[1] - 1 + j, u0[2] - 1 + k
                    s += arr[0, w, v, u] * keys(p[0] - u - 0.5) * keys(p[1] - v - 0.5) * keys(p[2] - w - 0.5)
        assert abs(oracle.interp(arr, 1, 0, p[None], f64=True)[0] - s) < 1e-12


# ------------------------------------------------------------------ gradient
def test_gradient_ramp_closed_form():
    a = 3.0
    shape = (1, 7, 8, 9)
    ramp = _field(lambda x, y, z, t: a * (x - 0.5), shape).astype(np.float32)
    gx, gy, gz = oracle.gradient(ramp, conf(GradientSmooth=1))
    assert np.all(gx[0, :, :, 1:-1] == 18 * a)      # 9 (j,k) pairs x 2a
    assert np.all(gx[0, :, :, 0] == 9 * a) and np.all(gx[0, :, :, -1] == 9 * a)  # clamp-to-edge
    assert np.all(gy == 0) and np.all(gz == 0)
    gx0, gy0, gz0 = oracle.gradient(ramp, conf(GradientSmooth=0))
    assert np.all(gx0[0, :, :, 1:-1] == 2 * a) and np.all(gx0[0, :, :, 0] == a)
    assert np.all(gy0 == 0) and np.all(gz0 == 0)
    # roles of the ranges: XRange widens Gy/Gz, MainRange the derivative axis
    gx2, gy2, gz2 = oracle.gradient(ramp, conf(GradientSmooth=1, MainRange=2, XRange=0, YRange=0, ZRange=0))
    assert np.all(gx2[0, :, :, 2:-2] == a * (1 + 2 + 1 + 2))  # (p[x+1]-p[x-1]) + (p[x+2]-p[x-2]) = 2a + 4a


# ------------------------------------------------------------------ ring
@pytest.mark.parametrize("axes", [1, 2, 6])
def test_ring_geometry(axes):
    n = 2 * axes
    R = 0.5
    r32 = oracle.build_ring(axes, R)
    r64 = oracle.build_ring(axes, R, f64=True)
    assert r32.shape == (n * n, 3)
    np.testing.assert_allclose(np.linalg.norm(r64, axis=1), R, rtol=1e-12)
    np.testing.assert_allclose(r32, r64, atol=2e-7)
    # theta = 0 row: n identical points at the pole (0, 0, R)
    np.testing.assert_allclose(r64[:n], np.tile([0, 0, R], (n, 1)), atol=1e-15)
    # closed form of the angle grid (pi as the reference's 3.141592654)
    step = 2 * 3.141592654 / n
    i1, i2 = np.divmod(np.arange(n * n), n)
    np.testing.assert_allclose(r64[:, 2], R * np.cos(step * i1), atol=1e-15)
    np.testing.assert_allclose(r64[:, 0], R * np.sin(step * i1) * np.cos(step * i2), atol=1e-15)
    # product library builds the same FP32 ring
    lib_ring = srrf.ring(axes, R).xyz
    assert np.array_equal(lib_ring, r32)


def test_direction_sets():
    for kind, n in ((6, 6), (26, 26)):
        s = srrf.directions(kind, 0.5)
        assert s.n == n and s.divisor == n
        np.testing.assert_allclose(np.linalg.norm(s.xyz.astype(np.float64), axis=1), 0.5, rtol=1e-7)
        # symmetric set: sum of offsets is zero
        np.testing.assert_allclose(s.xyz.sum(0), 0, atol=1e-6)
        # grouped: equal (x, y) offsets are consecutive
        seen, prev = set(), None
        for p in map(tuple, s.xyz[:, :2]):
            if p != prev:
                assert p not in seen
                seen.add(p)
            prev = p


# ------------------------------------------------------------------ radiality
def small_series(shape=(4, 8, 10, 12), seed=0):
    from rfba_amd import synth
    T, Z, Y, X = shape
    return synth.srrf_series((X, Y, Z, T), 3, seed, threads=4)


@pytest.mark.parametrize("dirs", ["ring2", "d6", "d26"])
def test_radiality_exact_invariants(dirs):
    s = small_series()
    smp = srrf.ring(2, 0.5) if dirs == "ring2" else srrf.directions(int(dirs[1:]), 0.5)
    c = conf(Method=1, IntensityW=0)
    base = oracle.srrf_stack(s.astype(np.float32), c, smp.xyz, smp.divisor)
    # power-of-two scaling commutes with every rounding
    scaled = oracle.srrf_stack(s.astype(np.float32) * 4, c, smp.xyz, smp.divisor)
    assert np.array_equal(base, scaled, equal_nan=True)
    # integer offset: integer-valued gradients are unchanged exactly
    shifted = oracle.srrf_stack(s.astype(np.float32) + 1000, c, smp.xyz, smp.divisor)
    assert np.array_equal(base, shifted, equal_nan=True)
    # NaN band: interpolation returns 0 near faces -> NaN radiality
    assert np.isnan(base).any() and np.isfinite(base).any()


def test_constant_input_is_nan():
    c = conf(Method=1, IntensityW=0)
    smp = srrf.directions(6, 0.5)
    out = oracle.srrf_stack(np.full((2, 8, 8, 8), 7, np.float32), c, smp.xyz, smp.divisor)
    assert np.isnan(out).all()
    out0 = oracle.srrf_stack(np.full((2, 8, 8, 8), 7, np.float32), conf(Method=1, RemoveCon=1), smp.xyz,
                             smp.divisor)
    assert np.all(out0 == 0)  # fmaxf(NaN, 0) = 0 under the positivity constraint


def test_radial_blob_converges():
    # Gaussian blob centred on an SR voxel centre: every sample converges and
    # the gradient lines pass (nearly) through the centre -> radiality ~ 1.
    M = 2
    shape = (1, 12, 12, 12)
    cx = 6.0  # low-res coordinate of SR voxel centre X=11 -> (11+0.5)/2 = 5.75
    Xs = 11
    c0 = (Xs + 0.5) / M
    blob = _field(lambda x, y, z, t: 1000 * np.exp(-((x - c0) ** 2 + (y - c0) ** 2 + (z - c0) ** 2) / (2 * 1.5 ** 2)),
                  shape).astype(np.float32)
    for smp in (srrf.directions(6, 0.5), srrf.directions(26, 0.5)):
        out = oracle.srrf_stack(blob, conf(Method=1, IntensityW=0, GradientSmooth=0, Magnification=M), smp.xyz,
                                smp.divisor)
        assert out[Xs, Xs, Xs] > 0.5  # all samples converge; sampled-Gaussian gradients are not exactly radial
        fin = np.nan_to_num(out, nan=-np.inf)
        assert np.unravel_index(np.argmax(fin), out.shape) == (Xs, Xs, Xs)


def test_fp32_vs_fp64_close():
    s = small_series(seed=3).astype(np.float32)
    smp = srrf.directions(26, 0.5)
    c = conf(Method=1, IntensityW=1, Magnification=3)
    o32 = oracle.srrf_stack(s, c, smp.xyz, smp.divisor)
    o64 = oracle.srrf_stack(s.astype(np.float64), c, smp.xyz.astype(np.float64), smp.divisor, f64=True)
    assert np.array_equal(np.isnan(o32), np.isnan(o64))
    m = np.isfinite(o64)
    err = np.abs(o32[m] - o64[m]).max() / np.abs(o64[m]).max()
    assert err < 1e-4


# ------------------------------------------------------------------ temporal closed forms
def temporal_of(series_1d, method, itc=0, order=2, iw=0):
    T = len(series_1d)
    rad = np.asarray(series_1d, np.float32).reshape(T, 1, 1, 1)
    img = np.zeros((T, 1, 1, 1), np.float32)
    c = conf(Method=method, ITC=itc, TRACOrder=order, IntensityW=iw, Magnification=1)
    return float(oracle.temporal(img, rad, c)[0, 0, 0])


def test_temporal_hand_derived():
    assert temporal_of([0.25] * 8, 1) == 0.25                        # TRA mean
    assert temporal_of([-1.0, -2.0], 0) == 0.0                       # TRM starts from 0
    assert temporal_of([0.5, 3.0, -1.0], 0) == 3.0
    assert temporal_of([1.0, -1.0, 2.0], 2) == pytest.approx(7 / 6, rel=1e-7)  # TRPPM counts skipped pairs
    assert temporal_of([0.75] * 5, 2) == 0.5625                      # TRPPM constant c -> c^2
    alt = [1.0, -1.0] * 4
    assert temporal_of(alt, 3, itc=0) == 6 / 8                        # |sum_{t<T-2}(r_t r_t+1)|/T
    # ITC order 2, r = 1..8: level 0 AB = 17.5 -> 17.5/8; binning zeroes r[0]
    # before reading it, so bins are (1, 3.5, 5.5, 7.5); level 1 AB = 2.5 -> 2.5/4
    assert temporal_of(np.arange(1, 9), 3, itc=1) == 2.1875 + 0.625
    # non-ITC order 3 and 4 on r = 1..8 (mean 4.5), direct sums
    r = np.arange(1, 9) - 4.5
    abc = sum(r[t] * r[t + 1] * r[t + 2] for t in range(8 - 3))
    assert temporal_of(np.arange(1, 9), 3, itc=0, order=3) == pytest.approx(abs(abc) / 8, rel=1e-6)
    t4 = range(8 - 4)
    A, B, C, D = (np.array([r[t + k] for t in t4]) for k in range(4))
    v4 = (A * B * C * D).sum() - (A * B).sum() * (C * D).sum() - (A * C).sum() * (B * D).sum() - (A * D).sum() * (B * C).sum()
    assert temporal_of(np.arange(1, 9), 3, itc=0, order=4) == pytest.approx(abs(v4) / 8, rel=1e-6)


def test_temporal_itc_odd_length_reads_zero_past_end():
    # T = 5, order 2: the level-0 binning reads r[5]; defined as 0.
    assert temporal_of([1.0, 2.0, 3.0, 4.0, 5.0], 3, itc=1) == pytest.approx(0.4, rel=1e-7)


def test_temporal_intensity_weighting_constant_image():
    T = 6
    rad = np.full((T, 2, 2, 2), 0.5, np.float32)
    img = np.full((T, 1, 1, 1), 8.0, np.float32)
    # 1x1x1 low-res: every centre is outside [1.5, dim-1.5] -> I = 0
    c = conf(Method=1, IntensityW=1, Magnification=2)
    assert np.all(oracle.temporal(img, rad, c) == 0)
    img2 = np.full((T, 6, 6, 6), 8.0, np.float32)
    rad2 = np.full((T, 12, 12, 12), 0.5, np.float32)
    out = oracle.temporal(img2, rad2, c)
    # interior voxel: constant image interpolates to 8 (partition of unity)
    assert out[6, 6, 6] == pytest.approx(4.0, rel=1e-6)


# ------------------------------------------------------------------ tiling
def test_tiled_single_tile_equals_stack_and_leaves_low_band():
    s = small_series((3, 10, 11, 12), seed=5).astype(np.float32)
    c = conf(Method=1, IntensityW=1, AxesinRing=2, Magnification=2, blockSizeX=16, blockSizeY=16, blockSizeZ=16)
    tiled = oracle.srrf_tiled(s, c)
    smp = oracle.build_ring(2, 0.5)
    full = oracle.srrf_stack(s, c, smp, 4.0)
    M = 2
    assert np.isnan(tiled[:2 * M]).all() and np.isnan(tiled[:, :2 * M]).all() and np.isnan(tiled[:, :, :2 * M]).all()
    core = (slice(2 * M, None),) * 3
    assert np.array_equal(tiled[core], full[core], equal_nan=True)


def test_tiled_split_counts():
    # getSplitCount: ceil((dim-4)/(B-4)); dim 4 -> no tile at all on that axis
    s = np.zeros((1, 4, 20, 20), np.float32)
    c = conf(Magnification=1, AxesinRing=1)
    out = oracle.srrf_tiled(s, c, fill=-5.0)
    assert np.all(out == -5.0)
