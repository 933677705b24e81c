"""Block geometry of RFBA's per-block runs and the batched workloads.

do_split restates the MATLAB block planner (Matlab Scripts/doSplit.m:1-20):
1-based [start, length] pairs along one axis, blocks of maxX voxels whose
neighbours overlap by 2*border, the last block taking the remainder.
interior_pixels restates the per-block result filter of loadResult.m:33-49.
"""
from __future__ import annotations

import numpy as np


def do_split(xall: int, maxX: int, border: int):
    """doSplit.m:3-16 -> list of (start, length), start 1-based."""
    if maxX >= xall:
        return [(1, xall)]
    pos = [[1, maxX]]
    while pos[-1][0] + maxX <= xall:
        pos.append([pos[-1][0] + pos[-1][1] - 2 * border, maxX])
    pos[-1][1] = xall - pos[-1][0] + 1
    return [tuple(p) for p in pos]


def volume_blocks(shape_xyz, block=29, border=3):
    """All blocks of a volume as (origin_xyz 0-based, dims_xyz), x fastest."""
    sx, sy, sz = (do_split(n, block, border) for n in shape_xyz)
    out = []
    for (z0, lz) in sz:
        for (y0, ly) in sy:
            for (x0, lx) in sx:
                out.append(((x0 - 1, y0 - 1, z0 - 1), (lx, ly, lz)))
    return out


def interior_pixels(pixels, border, x_list, y_list, z_list, block_id):
    """loadResult.m:33-49: keep the result columns [I, sigma, x, y, z]
    (1-based volume coordinates) of block block_id = (ix, iy, iz) whose
    block-local coordinate v - start + 1 lies strictly between border and
    length - border on every axis.  x_list etc. are do_split's (start, length)
    pairs; pixels is [5, n]."""
    pix = np.asarray(pixels, dtype=np.float64).reshape(5, -1)
    (sx, lx), (sy, ly), (sz, lz) = x_list[block_id[0]], y_list[block_id[1]], z_list[block_id[2]]
    x = pix[2] - sx + 1
    y = pix[3] - sy + 1
    z = pix[4] - sz + 1
    keep = (x > border) & (x < lx - border) & (y > border) & (y < ly - border) & (z > border) & (z < lz - border)
    return pix[:, keep]
