"""Block geometry of RFBA's per-block runs and the batched workloads.

do_split restates the MATLAB block planner (Matlab Scripts/doSplit.m:1-20):
1-based [start, length] pairs along one axis, blocks of maxX voxels whose
neighbours overlap by 2*border, the last block taking the remainder.
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
