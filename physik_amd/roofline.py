"""Algorithmic byte counts of the step's HBM-bound kernels.

These are the numerators of the roofline fractions bench.py reports and of
the counter summaries under profiles/ (physik_amd.profsum): the bytes a
kernel must move at the least, independent of the storage format, set
against the measured time (achieved GB/s) or the HBM counters (traffic).
"""

from __future__ import annotations


def spmv_bytes(blocks, rows, body_launches, s, upper_blocks=None):
    """Algorithmic bytes of SpMV work q = A d, format-independent: every
    stored 3x3 block the launches multiplied once (9 values of s bytes + an
    int32 column), per row one int32 row pointer, d read once and q written
    once (3 values each), and the closing row pointer of each body in each
    launch: nnzb (9 s + 4) + (N_d + 1) 4 + 2 N_d 3 s for one launch over a
    whole body. With the upper triangle stored (upper_blocks = the stored
    blocks among `blocks`), nnzb is the upper blocks: each also serves its
    mirror's transposed product."""
    stored = blocks if upper_blocks is None else upper_blocks
    return stored * (9 * s + 4) + rows * (4 + 2 * 3 * s) + 4 * body_launches


def spmv_format_bytes(blocks, rows, body_launches, s, upper_blocks=None):
    """Bytes the executed format itself holds for the same work: full
    storage as spmv_bytes; sliced upper-only storage holds a record of 10
    words of s bytes per upper block (the FP64 column word is 8 bytes) and an
    8-byte (mirror element, column) entry per lower block, and no row
    pointers (its slice descriptors and lane masks, 16 bytes and 8 bytes per
    ELL column for 64 rows, about 2 bytes per row, are not counted)."""
    if upper_blocks is None:
        return spmv_bytes(blocks, rows, body_launches, s, None)
    return upper_blocks * 10 * s + (blocks - upper_blocks) * 8 + rows * 2 * 3 * s


def assembly_bytes(tets, nodes, stored_blocks, rows, s_matrix):
    """Algorithmic bytes of one assembly (element + slot gather + right-hand
    side) in the executed types: T (16 + 22 s_p) for 4 int32 node ids and 22
    per-tet rest values (Dm^-1, rest edges, V, lambda, mu, c) in FP64,
    N 6 s_x for x and v read once (FP64), nnzb 9 s_A for A written once (the
    stored blocks: the upper triangle under upper storage; s_A the storage
    type) and N_d 3 s_b for b written once (FP64)."""
    return tets * (16 + 22 * 8) + nodes * 6 * 8 + stored_blocks * 9 * s_matrix + rows * 3 * 8


def precision_sizes(precision):
    """(bytes per stored matrix value, upper triangle stored) of a pk_precision."""
    from . import native as N
    s = 8 if precision in (N.PK_FP64, N.PK_FP64_UPPER) else 4
    return s, precision in (N.PK_FP32_UPPER, N.PK_FP64_UPPER)
