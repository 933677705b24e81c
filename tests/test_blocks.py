"""Block planner (rfba_amd.blocks.do_split) against hand-derived cases of
Matlab Scripts/doSplit.m: [start, length] pairs, 1-based, neighbours
overlapping by 2*border, the last block taking the remainder."""
import numpy as np

from rfba_amd.blocks import do_split, interior_pixels, volume_blocks


def test_single_block_when_it_fits():
    assert do_split(20, 29, 3) == [(1, 20)]
    assert do_split(29, 29, 3) == [(1, 29)]


def test_hand_cases():
    # 30 voxels, 29-wide blocks, border 3: 1 + 29 > 30 stops after one step
    # pos = [1 29; 24 29] -> last length 30 - 24 + 1 = 7
    assert do_split(30, 29, 3) == [(1, 29), (24, 7)]
    # xall = 10, maxX = 4, border 1: starts 1, 3, 5, 7 (step 2); 7 + 4 > 10
    assert do_split(10, 4, 1) == [(1, 4), (3, 4), (5, 4), (7, 4)]
    # border 0: plain tiling with a short last block
    assert do_split(10, 4, 0) == [(1, 4), (5, 4), (9, 2)]


def test_rfba_volume_split():
    x = do_split(512, 29, 3)
    assert len(x) == 22 and all(n == 29 for _, n in x)
    assert x[-1] == (484, 29) and x[-1][0] + x[-1][1] - 1 == 512
    z = do_split(128, 29, 3)
    assert z == [(1, 29), (24, 29), (47, 29), (70, 29), (93, 29), (116, 13)]
    blocks = volume_blocks((512, 512, 128), 29, 3)
    assert len(blocks) == 22 * 22 * 6
    assert blocks[0] == ((0, 0, 0), (29, 29, 29))
    assert blocks[-1] == ((483, 483, 115), (29, 29, 13))


def test_coverage_and_overlap():
    for xall, m, b in [(512, 29, 3), (128, 29, 3), (100, 16, 2), (57, 9, 4)]:
        parts = do_split(xall, m, b)
        covered = set()
        for s, n in parts:
            covered.update(range(s, s + n))
        assert covered == set(range(1, xall + 1))
        for (s0, n0), (s1, _) in zip(parts, parts[1:]):
            assert s0 + n0 - s1 == 2 * b


def test_interior_filter_hand_case():
    # loadResult.m: block (2nd x block starting at 24, length 29), border 3:
    # local x = X - 24 + 1 must satisfy 3 < x < 26, i.e. X in 27..48
    xl = do_split(512, 29, 3)
    yl = zl = [(1, 29)]
    X = np.array([26, 27, 48, 49, 30, 30], np.float64)
    Y = np.array([10, 10, 10, 10, 3, 4], np.float64)
    Z = np.full(6, 10.0)
    pix = np.stack([np.ones(6), np.ones(6), X, Y, Z])
    kept = interior_pixels(pix, 3, xl, yl, zl, (1, 0, 0))
    assert kept[2].tolist() == [27, 48, 30]
    assert kept[3].tolist() == [10, 10, 4]
