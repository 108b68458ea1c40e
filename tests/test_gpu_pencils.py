"""Pencil SpMV tiles (pk_system_set_spmv_pencils): each tile of a layered body
pairs a run of rows with the run a layer later. Only the grouping of the d.q
partial sums changes, so states agree with plain tiles to FP64 rounding
(relative 1e-12) and CG iteration counts within 1; upper-triangle and full
storage remain bit-identical to each other with pencils on.
"""

import numpy as np
import pytest

from physik_amd import native as N
from physik_amd import workloads as W

from test_gpu_staging import mixed_batch
from test_gpu_upper import twisted

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("case", ["twisted", "batch"])
@pytest.mark.parametrize("shape", [(8, 128, 512, 2), (4, 64, 256, 1)])
def test_pencil_tiles_match_plain_tiles(case, shape):
    w = twisted(n=(16, 6, 5)) if case == "twisted" else mixed_batch(bar=False)
    plain = W.make_system(w, N.PK_FP64_UPPER)
    pen = W.make_system(w, N.PK_FP64_UPPER)
    pen_full = W.make_system(w, N.PK_FP64)
    for s in (plain, pen, pen_full):
        s.tune(*shape)
    pen.set_pencils(True)
    pen_full.set_pencils(True)
    for n in range(w.steps):
        for s in (plain, pen, pen_full):
            s.step(w.dt)
        xp, vp = plain.state()
        x, v = pen.state()
        xf, vf = pen_full.state()
        print(f"{case} {shape} step {n}: rel x {rel(x, xp):.2e} v {rel(v, vp):.2e}", flush=True)
        assert rel(x, xp) <= 1e-12 and rel(v, vp) <= 1e-12
        assert np.abs(pen.results()[0] - plain.results()[0]).max() <= 1
        assert np.array_equal(x, xf) and np.array_equal(v, vf), "upper and full storage with pencils"
    for s in (plain, pen, pen_full):
        s.close()
