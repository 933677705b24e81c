"""GPU parity of the fused SRRF kernel (librfba_srrf.so) against the FP32
oracle (oracle/srrf_oracle.cpp) on the same seeded inputs.

Bar: NaN masks identical and |gpu - oracle| <= 1e-5 * |oracle| + 1e-7 * max|oracle|
elementwise (the kernel reproduces the reference arithmetic operation by
operation, so in practice the outputs are bit-identical; the number of
differing voxels is printed).  The FP64 re-evaluation error of both is
printed per case.
"""
import os

import numpy as np
import pytest

from oracle import oracle
from rfba_amd import srrf, synth

pytestmark = pytest.mark.gpu

RTOL = 1e-5
ATOL_FRAC = 1e-7


def conf(**kw):
    c = srrf.SRRFConf()
    base = dict(RingRadius=0.5, Magnification=2, AxesinRing=2, MainRange=1, XRange=1, YRange=1, ZRange=1,
                Method=1, ITC=0, TRACOrder=2, RemoveCon=0, Renormalize=0, GradientSmooth=1, IntensityW=0,
                blockSizeX=16, blockSizeY=16, blockSizeZ=16)
    base.update(kw)
    for k, v in base.items():
        setattr(c, k, v)
    return c


def parity(gpu, ref, label=""):
    assert gpu.shape == ref.shape
    nan_g, nan_r = np.isnan(gpu), np.isnan(ref)
    assert np.array_equal(nan_g, nan_r), "%s: NaN masks differ (%d vs %d)" % (label, nan_g.sum(), nan_r.sum())
    m = ~nan_r
    if not m.any():
        return 0.0
    g, r = gpu[m].astype(np.float64), ref[m].astype(np.float64)
    scale = np.abs(r).max()
    diff = np.abs(g - r)
    ndiff = int((g != r).sum())
    tol = RTOL * np.abs(r) + ATOL_FRAC * scale
    worst = float((diff / (np.abs(r) + ATOL_FRAC / RTOL * scale + 1e-300)).max())
    print("%s: %d/%d voxels differ, max |d| %.3g, scale %.3g" % (label, ndiff, m.sum(), diff.max(), scale))
    assert np.all(diff <= tol), "%s: max diff %g" % (label, diff.max())
    return worst


def series(shape_tzyx, n_em, seed):
    T, Z, Y, X = shape_tzyx
    return synth.srrf_series((X, Y, Z, T), n_em, seed)


def test_config1_six_directions_tra(gpu):
    # 48x48x24, T=50, 12 emitters, M=2, R=0.5, 6 axis directions, TRA
    s = series((50, 24, 48, 48), 12, 1)
    smp = srrf.directions(6, 0.5)
    for iw in (0, 1):
        c = conf(Method=1, IntensityW=iw, Magnification=2)
        g = srrf.srrf_stack(s, c, smp)
        o = oracle.srrf_stack(s.astype(np.float32), c, smp.xyz, smp.divisor)
        parity(g, o, "C1 6dir TRA iw=%d" % iw)
        o64 = oracle.srrf_stack(s.astype(np.float64), c, smp.xyz.astype(np.float64), smp.divisor, f64=True)
        m = np.isfinite(o64)
        print("   err vs FP64: gpu %.3g  oracle %.3g (max-normalised)" % (
            np.abs(g[m] - o64[m]).max() / np.abs(o64[m]).max(), np.abs(o[m] - o64[m]).max() / np.abs(o64[m]).max()))


def test_config1_reference_ring(gpu):
    s = series((12, 24, 48, 48), 12, 1)
    c = conf(Method=1, IntensityW=0, Magnification=2, AxesinRing=2)
    g = srrf.srrf_stack(s, c, None)
    ring = oracle.build_ring(2, 0.5)
    o = oracle.srrf_stack(s.astype(np.float32), c, ring, 4.0)
    parity(g, o, "C1 ring2 TRA")


@pytest.mark.parametrize("method,itc,order,iw", [
    (0, 0, 2, 0), (0, 0, 2, 1), (1, 0, 2, 1), (2, 0, 2, 0), (2, 0, 2, 1),
    (3, 0, 2, 0), (3, 0, 3, 1), (3, 0, 4, 0), (3, 1, 2, 1), (3, 1, 3, 0), (3, 1, 4, 1)])
def test_temporal_modes(gpu, method, itc, order, iw):
    s = series((9, 10, 18, 20), 4, 7)   # odd T exercises the ITC read past the end
    smp = srrf.directions(26, 0.5)
    c = conf(Method=method, ITC=itc, TRACOrder=order, IntensityW=iw, Magnification=2)
    g = srrf.srrf_stack(s, c, smp)
    o = oracle.srrf_stack(s.astype(np.float32), c, smp.xyz, smp.divisor)
    parity(g, o, "method %d itc %d ord %d iw %d" % (method, itc, order, iw))


@pytest.mark.parametrize("dtype", ["u16", "f32"])
@pytest.mark.parametrize("method,itc,order", [(0, 0, 2), (1, 0, 2), (2, 0, 2), (3, 0, 2), (3, 1, 3)])
@pytest.mark.parametrize("poscon", [0, 1])
def test_idle_tiles(gpu, dtype, method, itc, order, poscon):
    # A tile none of whose voxels has its samples in range loads no halo and
    # forms no centre intensities where that cannot change an output (integer
    # input, or no intensity weighting).  Two stacks: one 3 voxels deep
    # (every voxel NaN by definition, every tile idle) and one 29 wide at
    # M = 3, whose tiles the planner shifts so that a column of idle tiles
    # lies beside the computed ones.  Outputs equal the oracle's bit for bit
    # (the sign of zero included).
    smp = srrf.directions(26, 0.5)
    for iw in ((0, 1) if dtype == "u16" else (0,)):
        for shp, M in (((5, 3, 9, 11), 2), ((5, 6, 10, 29), 3)):
            s = series(shp, 3, 17)
            if dtype == "f32":
                s = s.astype(np.float32)
            c = conf(Method=method, ITC=itc, TRACOrder=order, IntensityW=iw, RemoveCon=poscon, Magnification=M)
            g = srrf.srrf_stack(s, c, smp)
            o = oracle.srrf_stack(s.astype(np.float32), c, smp.xyz, smp.divisor)
            lab = "%s method %d itc %d order %d iw %d poscon %d shape %s M %d" % (
                dtype, method, itc, order, iw, poscon, shp, M)
            assert np.array_equal(np.isnan(g), np.isnan(o)), lab
            m = ~np.isnan(o)
            assert np.array_equal(g[m].view(np.uint32), o[m].view(np.uint32)), lab


@pytest.mark.parametrize("mag,dirs", [(1, "d6"), (3, "d26"), (4, "d26"), (5, "ring6"), (2, "ring3")])
def test_magnifications_and_sets(gpu, mag, dirs):
    s = series((4, 9, 13, 15), 4, 11)
    if dirs.startswith("ring"):
        axes = int(dirs[4:])
        smp = None
        xyz, div = oracle.build_ring(axes, 0.5), 2.0 * axes
    else:
        smp = srrf.directions(int(dirs[1:]), 0.5)
        xyz, div = smp.xyz, smp.divisor
    c = conf(Method=1, IntensityW=1, Magnification=mag, AxesinRing=int(dirs[4:]) if smp is None else 2)
    g = srrf.srrf_stack(s, c, smp)
    o = oracle.srrf_stack(s.astype(np.float32), c, xyz, div)
    parity(g, o, "M=%d %s" % (mag, dirs))


@pytest.mark.parametrize("smooth,renorm,poscon,ranges", [
    (0, 0, 0, (1, 0, 0, 0)), (1, 1, 0, (1, 1, 1, 1)), (1, 0, 1, (2, 0, 1, 0)), (0, 1, 1, (1, 0, 0, 0))])
def test_gradient_and_radiality_options(gpu, smooth, renorm, poscon, ranges):
    s = series((3, 11, 12, 14), 3, 13)
    smp = srrf.directions(26, 0.5)
    mR, xR, yR, zR = ranges
    c = conf(Method=1, IntensityW=1, GradientSmooth=smooth, Renormalize=renorm, RemoveCon=poscon,
             MainRange=mR, XRange=xR, YRange=yR, ZRange=zR, Magnification=3)
    g = srrf.srrf_stack(s, c, smp)
    o = oracle.srrf_stack(s.astype(np.float32), c, smp.xyz, smp.divisor)
    parity(g, o, "smooth %d renorm %d poscon %d ranges %s" % (smooth, renorm, poscon, ranges))


@pytest.mark.parametrize("ranges", [(1, 1, 1, 1), (2, 3, 3, 3), (3, 3, 3, 3)])
def test_integer_input_gradient_sums_exact(gpu, ranges):
    # uint16 frames near full scale.  With at most 256 terms per smoothed
    # gradient sum (the first two range sets: 12 and 196 terms) the kernel
    # forms the gradient from shared box sums, exact for integers below
    # 2^24; with more (294 terms) it sums term by term.  Either way the
    # result equals the float-input path (term by term) and the oracle bit
    # for bit.
    rng = np.random.default_rng(19)
    s = rng.integers(50000, 65536, (3, 12, 13, 14)).astype(np.uint16)
    smp = srrf.directions(26, 0.5)
    mR, xR, yR, zR = ranges
    c = conf(Method=1, IntensityW=1, GradientSmooth=1, MainRange=mR, XRange=xR, YRange=yR, ZRange=zR,
             Magnification=2)
    g = srrf.srrf_stack(s, c, smp)
    gf = srrf.srrf_stack(s.astype(np.float32), c, smp)
    o = oracle.srrf_stack(s.astype(np.float32), c, smp.xyz, smp.divisor)
    print("ranges %s: %d finite voxels" % (ranges, int(np.isfinite(o).sum())))
    assert np.isfinite(o).any()
    assert np.array_equal(g, gf, equal_nan=True)
    assert np.array_equal(g, o, equal_nan=True)


def test_float_input_and_drop_in_gpu_entry(gpu):
    # normalised float frames, as the reference caller passes them
    s = series((5, 10, 14, 16), 4, 17).astype(np.float32)
    s = (s - s.mean()) / s.std()
    c = conf(Method=3, ITC=1, TRACOrder=2, IntensityW=1, Magnification=2, AxesinRing=2)
    g = srrf.SRRFGradientCalcGPU(s, c)
    o = oracle.srrf_stack(s, c, oracle.build_ring(2, 0.5), 4.0)
    parity(g, o, "SRRFGradientCalcGPU float TRAC ITC")


def test_drop_in_tiled_entry(gpu, tmp_path, monkeypatch):
    # SRRFGradientCalc reads ./SRRF.ini and tiles 16^3 with border 2
    (tmp_path / "SRRF.ini").write_text(
        "[SRRF]\nRingRadius=0.5\nMagnification=2\nAxesinRing=2\nMainRange=1\nXRange=1\nYRange=1\nZRange=1\n"
        "[Temproal]\nMethod=1\n[Radiality]\nGradient_Smoothing=1\n[Weighting]\nIntensity_Weighting=1\n")
    monkeypatch.chdir(tmp_path)
    s = series((4, 20, 29, 33), 6, 19).astype(np.float32)
    g = srrf.SRRFGradientCalc(s, fill=np.nan)
    c = srrf.load_ini("SRRF.ini")
    o = oracle.srrf_tiled(s, c, fill=np.nan)
    parity(g, o, "SRRFGradientCalc tiled")


def test_batched_matches_single_stacks(gpu):
    rng = np.random.default_rng(5)
    src = series((11, 14, 30, 33), 10, 23)
    d_src = srrf.DeviceBuffer.from_array(src)
    smp = srrf.directions(6, 0.5)
    c = conf(Method=3, ITC=1, IntensityW=1)
    blocks, refs = [], []
    for i in range(6):
        dims = tuple(int(v) for v in (rng.integers(6, 16), rng.integers(6, 14), rng.integers(5, 9)))
        org = tuple(int(rng.integers(0, lim - d + 1)) for lim, d in zip((33, 30, 14), dims))
        frames = int(rng.integers(3, 12))
        mag = int(rng.integers(2, 5))
        out = srrf.DeviceBuffer(dims[0] * dims[1] * dims[2] * mag ** 3 * 4)
        blocks.append(srrf.Block(d_src, src.shape, org, dims, frames, mag, out))
        sub = src[:frames, org[2]:org[2] + dims[2], org[1]:org[1] + dims[1], org[0]:org[0] + dims[0]]
        refs.append(srrf.srrf_stack(np.ascontiguousarray(sub), c.replace(Magnification=mag), smp))
    b = srrf.Batch(blocks, srrf.RFBA_U16, c, smp)
    for mode in ("graph", "plain", "per_block", "graph", "async"):
        for bl in blocks:
            bl.out.zero()
        if mode == "per_block":
            b.run_per_block()
        elif mode == "async":
            b.run_async(use_graph=True)
            b.status()
        else:
            b.run(use_graph=(mode == "graph"))
        for i, bl in enumerate(blocks):
            X, Y, Z = bl.dims_xyz
            M = bl.mag
            got = bl.out.to_array((Z * M, Y * M, X * M), np.float32)
            assert np.array_equal(got, refs[i], equal_nan=True), (mode, i)
    b.close()


def test_autotuned_batch_matches_single_stacks(gpu):
    # uneven blocks and frame counts (TRAC keeps a per-voxel series in
    # scratch, whose size depends on the tile depth): after autotuning, and
    # at every tile depth the tuner times, results stay bit-identical.
    rng = np.random.default_rng(11)
    src = series((40, 12, 26, 28), 8, 31)
    d_src = srrf.DeviceBuffer.from_array(src)
    smp = srrf.directions(6, 0.5)
    c = conf(Method=3, ITC=0, TRACOrder=2, IntensityW=1)
    blocks, refs = [], []
    for i in range(7):
        dims = tuple(int(v) for v in (rng.integers(5, 14), rng.integers(5, 14), rng.integers(4, 12)))
        org = tuple(int(rng.integers(0, lim - d + 1)) for lim, d in zip((28, 26, 12), dims))
        frames = int(rng.integers(4, 41))
        mag = int(rng.integers(2, 5))
        out = srrf.DeviceBuffer(dims[0] * dims[1] * dims[2] * mag ** 3 * 4)
        blocks.append(srrf.Block(d_src, src.shape, org, dims, frames, mag, out))
        sub = src[:frames, org[2]:org[2] + dims[2], org[1]:org[1] + dims[1], org[0]:org[0] + dims[0]]
        refs.append(srrf.srrf_stack(np.ascontiguousarray(sub), c.replace(Magnification=mag), smp))
    b = srrf.Batch(blocks, srrf.RFBA_U16, c, smp)
    b.run(use_graph=True)  # a graph captured before tuning must be dropped
    tz, ms = b.autotune(sample_stride=2)
    print("autotune: depth %d, sample ms %s, groups %s" % (tz, ms, b.groups))
    assert tz in ms and ms[tz] == min(ms.values()) and b.tile_depth == tz
    assert set(ms) <= {4, 8, 16} and 8 in ms
    assert sorted(m for m, _ in b.groups) == sorted(set(bl.mag for bl in blocks))
    for mode in ("graph", "per_block"):
        for bl in blocks:
            bl.out.zero()
        if mode == "per_block":
            b.run_per_block()
        else:
            b.run(use_graph=True)
        for i, bl in enumerate(blocks):
            X, Y, Z = bl.dims_xyz
            M = bl.mag
            got = bl.out.to_array((Z * M, Y * M, X * M), np.float32)
            assert np.array_equal(got, refs[i], equal_nan=True), (mode, i, tz)
    b.close()
    d_src.close()


def test_full_size_boxes(gpu):
    # Full-width volume (256 x 256 low-res, M=4 -> 1024 x 1024 SR) against the
    # oracle evaluated on SR boxes with the same absolute coordinates: the
    # reference's float rounding depends on position, so parity must be
    # checked in place, not on a crop.
    s = series((12, 20, 256, 256), 150, 29)
    smp = srrf.directions(26, 0.5)
    c = conf(Method=3, ITC=0, TRACOrder=2, IntensityW=1, Magnification=4)
    full = srrf.srrf_stack(s, c, smp)
    for x0, y0, z0 in ((1000, 990, 40), (500, 611, 3), (0, 0, 0), (1008, 1008, 64)):
        o = oracle.srrf_stack_box(s.astype(np.float32), c, smp.xyz, smp.divisor, (x0, y0, z0), (16, 16, 16))
        parity(full[z0:z0 + 16, y0:y0 + 16, x0:x0 + 16], o, "full-size box at %s" % ((x0, y0, z0),))
    # power-of-two scaling of the input leaves radiality unchanged exactly;
    # with intensity weighting the output scales by the same power of two
    full2 = srrf.srrf_stack(s.astype(np.float32) * 2, c, smp)
    assert np.array_equal(full2, full * 2, equal_nan=True)


@pytest.mark.parametrize("radius", [0.6, 0.75, 1.0, 5e-7])
def test_radius_variants(gpu, radius):
    # non power-of-two radius (division path), a larger radius (wider halo)
    # and one below the fast path's range (every sample takes the general form)
    s = series((4, 10, 13, 15), 4, 31)
    smp = srrf.directions(26, radius)
    c = conf(Method=3, ITC=1, IntensityW=1, Magnification=3, RingRadius=radius)
    g = srrf.srrf_stack(s, c, smp)
    o = oracle.srrf_stack(s.astype(np.float32), c, smp.xyz, smp.divisor)
    parity(g, o, "R=%g" % radius)


@pytest.mark.parametrize("method,poscon", [(1, 0), (0, 0), (3, 0), (2, 1)])
def test_float_input_with_nan_and_inf(gpu, method, poscon):
    # non-finite input values propagate exactly as in the reference arithmetic
    # (gradient sums, interpolation, the general radiality term)
    s = series((9, 10, 14, 13), 6, 61).astype(np.float32)
    s[2, 5, 7, 6] = np.nan
    s[4, 3, 3, 9] = np.inf
    s[6, 7, 10, 2] = -np.inf
    smp = srrf.directions(26, 0.5)
    c = conf(Method=method, RemoveCon=poscon, IntensityW=1, Magnification=2, ITC=0, TRACOrder=2)
    g = srrf.srrf_stack(s, c, smp)
    o = oracle.srrf_stack(s, c, smp.xyz, smp.divisor)
    assert np.array_equal(g, o, equal_nan=True)
    print("method %d: %d of %d outputs NaN" % (method, np.isnan(o).sum(), o.size))


def _c3_blocks(src, specs):
    """Blocks (origin, dims, frames, mag) of a source series, with their
    device outputs."""
    d_src = srrf.DeviceBuffer.from_array(src)
    blocks = []
    for org, dims, frames, mag in specs:
        out = srrf.DeviceBuffer(dims[0] * dims[1] * dims[2] * mag ** 3 * 4)
        blocks.append(srrf.Block(d_src, src.shape, org, dims, frames, mag, out))
    return d_src, blocks


def _boxes(block):
    """Three 16^3 SR boxes of a block: low corner, centre, high corner."""
    X, Y, Z = (d * block.mag for d in block.dims_xyz)
    n = [min(16, v) for v in (X, Y, Z)]
    lo = (0, 0, 0)
    mid = tuple(max(0, (v - k) // 2) for v, k in zip((X, Y, Z), n))
    hi = tuple(v - k for v, k in zip((X, Y, Z), n))
    return [lo, mid, hi], n


@pytest.mark.parametrize("method", ["tra", "trac"])
def test_c3_shaped_batch_vs_oracle(gpu, method):
    # Blocks shaped like the batched workload: 29^3 doSplit blocks at M = 2,
    # 3, 4, a 29 x 29 x 13 z-edge block, an 8^3 extra stack and uneven ones,
    # frame counts 30-200; TRA with 6 axis directions and intensity
    # weighting (TRAC order 2 at T = 200 for the series scratch).  The
    # autotuned batched plan is compared, box by box, with the FP32 oracle
    # evaluated on each block as its own stack.
    src = series((200, 48, 60, 60), 40, 33)
    if method == "tra":
        specs = [((0, 0, 0), (29, 29, 29), 120, 2), ((26, 26, 0), (29, 29, 29), 30, 3),
                 ((31, 5, 19), (29, 29, 29), 200, 4), ((3, 29, 35), (29, 29, 13), 120, 3),
                 ((50, 50, 40), (8, 8, 8), 57, 2), ((40, 11, 2), (17, 23, 21), 163, 4)]
        c = conf(Method=1, IntensityW=1)
    else:
        specs = [((5, 7, 9), (29, 29, 29), 200, 2), ((30, 30, 19), (29, 29, 29), 200, 4),
                 ((50, 44, 40), (8, 9, 8), 200, 3)]
        c = conf(Method=3, ITC=0, TRACOrder=2, IntensityW=1)
    smp = srrf.directions(6, 0.5)
    d_src, blocks = _c3_blocks(src, specs)
    b = srrf.Batch(blocks, srrf.RFBA_U16, c, smp)
    tz, ms = b.autotune()
    print("%s: groups %s, tuned depth %d" % (method, b.groups, tz))
    # Every SR voxel of every block is written, by the batched and the
    # per-block form alike: outputs start as 0x7f bytes (3.4e38, a value no
    # output takes), and the two forms agree on the whole block.
    per_block = []
    for run in ("per_block", "batched"):
        for bl in blocks:
            bl.out.fill_byte(0x7F)
        if run == "per_block":
            b.run_per_block()
        else:
            b.run(use_graph=True)
        for bl in blocks:
            (dx, dy, dz), M = bl.dims_xyz, bl.mag
            got = bl.out.to_array((dz * M, dy * M, dx * M), np.float32)
            assert not np.any(got.view(np.uint32) == 0x7F7F7F7F), (run, bl.origin_xyz, bl.dims_xyz, M)
            if run == "per_block":
                per_block.append(got)
            else:
                assert np.array_equal(got, per_block.pop(0), equal_nan=True), (bl.origin_xyz, bl.dims_xyz, M)
    nbox = 0
    for bl in blocks:
        (ox, oy, oz), (dx, dy, dz) = bl.origin_xyz, bl.dims_xyz
        M = bl.mag
        got = bl.out.to_array((dz * M, dy * M, dx * M), np.float32)
        sub = np.ascontiguousarray(src[:bl.frames, oz:oz + dz, oy:oy + dy, ox:ox + dx]).astype(np.float32)
        starts, n = _boxes(bl)
        cm = c.replace(Magnification=M)
        for x0, y0, z0 in starts:
            o = oracle.srrf_stack_box(sub, cm, smp.xyz, smp.divisor, (x0, y0, z0), tuple(n))
            parity(got[z0:z0 + n[2], y0:y0 + n[1], x0:x0 + n[0]], o,
                   "%s block %s+%s T=%d M=%d box %s" % (method, bl.origin_xyz, bl.dims_xyz, bl.frames, M,
                                                         (x0, y0, z0)))
            nbox += 1
    b.close()
    d_src.close()
    assert nbox == 3 * len(blocks)


@pytest.mark.parametrize("method,ndirs", [(1, 6), (3, 26), (0, 26)])
def test_batch_unusual_geometries(gpu, method, ndirs):
    # Magnifications 1, 5 and 6 (tile phases at every residue of 16), single
    # frames, and blocks only 3 voxels wide or deep (every voxel NaN by
    # definition, every tile idle) in one batch: every SR voxel is written,
    # the batched and the per-block form agree bit for bit, and both equal
    # the oracle evaluated on each block as its own stack.
    src = series((7, 14, 15, 16), 6, 41)
    specs = [((0, 0, 0), (9, 10, 11), 7, 1), ((3, 2, 1), (7, 8, 9), 1, 5), ((5, 5, 5), (6, 5, 4), 2, 6),
             ((1, 1, 1), (3, 12, 10), 7, 3), ((2, 3, 4), (11, 9, 3), 5, 2), ((0, 4, 0), (16, 4, 14), 3, 4)]
    c = conf(Method=method, IntensityW=1, ITC=0, TRACOrder=2)
    smp = srrf.directions(ndirs, 0.5)
    d_src, blocks = _c3_blocks(src, specs)
    b = srrf.Batch(blocks, srrf.RFBA_U16, c, smp)
    outs = {}
    for run in ("per_block", "batched"):
        for bl in blocks:
            bl.out.fill_byte(0x7F)
        if run == "per_block":
            b.run_per_block()
        else:
            b.run(use_graph=True)
        outs[run] = []
        for bl in blocks:
            (dx, dy, dz), M = bl.dims_xyz, bl.mag
            got = bl.out.to_array((dz * M, dy * M, dx * M), np.float32)
            assert not np.any(got.view(np.uint32) == 0x7F7F7F7F), (run, bl.origin_xyz, bl.dims_xyz, M)
            outs[run].append(got)
    for i, bl in enumerate(blocks):
        (ox, oy, oz), (dx, dy, dz) = bl.origin_xyz, bl.dims_xyz
        lab = "block %s+%s T=%d M=%d" % (bl.origin_xyz, bl.dims_xyz, bl.frames, bl.mag)
        assert np.array_equal(outs["batched"][i], outs["per_block"][i], equal_nan=True), lab
        sub = np.ascontiguousarray(src[:bl.frames, oz:oz + dz, oy:oy + dy, ox:ox + dx]).astype(np.float32)
        o = oracle.srrf_stack(sub, c.replace(Magnification=bl.mag), smp.xyz, smp.divisor)
        assert np.array_equal(outs["batched"][i], o, equal_nan=True), lab
    b.close()
    d_src.close()
