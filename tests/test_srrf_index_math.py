"""CPU test: the SRRF kernel's division-free index decomposition.

srrf_fused decomposes the flat index e of its per-frame element loops with
q = (int)((float(e) + 0.5f) * r), r the correctly rounded FP32 1/d, instead
of integer division by a runtime extent d.  The claim: for 0 <= e < 2^16
(LDS extents keep e < 40960) and d >= 1 this is exactly e // d, because
(e + 0.5) / d lies at least 0.5 / d from an integer and the two FP32
roundings move it by less than 0.01 / d.  Checked exhaustively here in FP32
(round to nearest, no fused multiply-add, as the kernel is compiled) for
every e < 2^16 and every d <= 4096 (tile extents stay far below that).
"""
import numpy as np


def test_float_reciprocal_quotient_is_exact():
    e = np.arange(1 << 16, dtype=np.int64)
    ef = e.astype(np.float32) + np.float32(0.5)
    for d in range(1, 4097):
        r = np.float32(1.0) / np.float32(d)
        q = np.trunc(ef * r).astype(np.int64)
        assert np.array_equal(q, e // d), d
