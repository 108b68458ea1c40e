"""Launch shapes for tests that run one shape over several storages.

Upper-only storage (PK_FP32_UPPER, PK_FP64_UPPER) runs its SpMV with one lane
per row over whole 64-row slices; ``shape_for`` turns a full-storage shape
(lanes, spmv rows, vector rows, unroll) into the shape that storage accepts:
one lane, the rows rounded up to a multiple of 64, the rest unchanged. Full
storage at one lane per row gives the same results as upper-only storage at
the same rows and unroll.
"""

from physik_amd import native as N

UPPER = (N.PK_FP32_UPPER, N.PK_FP64_UPPER)


# a shape both storages accept, at which they agree bit for bit
PAIR_SHAPE = (1, 512, 1024, 1)


def shape_for(precision, lanes, rows, vec, unroll):
    if precision in UPPER:
        return 1, max(64, -(-rows // 64) * 64), vec, unroll
    return lanes, rows, vec, unroll
